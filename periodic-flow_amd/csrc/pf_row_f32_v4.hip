#define PF_REAL float
#define PF_VAR 4
#include "pf_row.inc"
