// Stage kernels with surface forcing, sponge relaxation and Chapman radiation edges (FEAT = 7).
#define RKG_FEAT 7
#include "stage_feat.inc"
