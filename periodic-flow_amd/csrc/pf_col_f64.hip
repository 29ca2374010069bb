#define PF_REAL double
#include "pf_col.inc"
