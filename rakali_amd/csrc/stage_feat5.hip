// Stage kernels with surface forcing and Chapman radiation edges (FEAT = 5).
#define RKG_FEAT 5
#include "stage_feat.inc"
