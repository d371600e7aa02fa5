#define IEQ_HALL_REAL float
#include "ieq_hall.inc"
