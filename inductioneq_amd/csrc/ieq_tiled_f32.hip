#define IEQ_TILED_REAL float
#include "ieq_tiled.inc"
