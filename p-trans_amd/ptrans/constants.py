"""Constants of the P-TRANS trace (src/RTMCphon.f90:62-79, 127-132) and of the C-ABI."""

M_PI = 3.14159265358979323846          # RT:62
M_TWOPI = M_PI * 2.0                   # RT:63
M_PI_2 = M_PI * 0.5                    # RT:64
HBAR = 1.054571e-34                    # RT:67
KB = 1.3806503e-23                     # RT:68
SMALL_VAL = 1.0e-5                     # RT:69
BALLISTIC_MFP = 1.0e10                 # RT:510, 1686

SHAPE_POLYGON, SHAPE_SPHERE, SHAPE_CYLINDER = 1, 2, 3        # RT:72-74
TRANS_VALUE, TRANS_GAMMA, TRANS_TBC = 1, 2, 3                # RT:77-79
TRANS_LORENTZ, TRANS_PERIODIC = 4, 5                         # extensions (include/ptrans.h)
MFP_MATTHIESSEN, MFP_SAMPLE = 1, 2                           # RT:127-128
GRAIN_VORONOI, GRAIN_BRICKS = 1, 2                           # RT:131-132
GRAIN_MESH = 3                                               # extension (include/ptrans.h)

# C-ABI status codes (include/ptrans.h)
PT_OK = 0
PT_E_INVALID = -1
PT_E_HIP = -2
PT_E_NOMEM = -3
PT_E_GEOMETRY = -4
PT_E_UNSUPPORTED = -5
PT_E_NODEVICE = -6

PREC_FP64 = 0
PREC_FP32_RETEST = 1
PREC_FP32_FILTER = 2

PATH_LOST = 1
PATH_START_MISMATCH = 2
PATH_EVENT_LIMIT = 4
