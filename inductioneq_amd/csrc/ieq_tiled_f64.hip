#define IEQ_TILED_REAL double
#define IEQ_TILED_DEFINE_OK
#include "ieq_tiled.inc"
