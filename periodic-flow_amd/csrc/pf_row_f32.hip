#define PF_REAL float
#include "pf_row.inc"
