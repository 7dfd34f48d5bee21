// Stage kernels with sponge relaxation (FEAT = 2).
#define RKG_FEAT 2
#include "stage_feat.inc"
