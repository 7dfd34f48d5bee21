// Stage kernels with Chapman radiation edges (FEAT = 4).
#define RKG_FEAT 4
#include "stage_feat.inc"
