#define IEQ_HALL_REAL double
#include "ieq_hall.inc"
