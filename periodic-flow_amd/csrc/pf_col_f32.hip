#define PF_REAL float
#include "pf_col.inc"
