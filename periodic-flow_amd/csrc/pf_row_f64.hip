#define PF_REAL double
#include "pf_row.inc"
