// Stage kernels with sponge relaxation and Chapman radiation edges (FEAT = 6).
#define RKG_FEAT 6
#include "stage_feat.inc"
