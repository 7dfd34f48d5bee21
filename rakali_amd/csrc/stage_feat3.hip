// Stage kernels with surface forcing and sponge relaxation (FEAT = 3).
#define RKG_FEAT 3
#include "stage_feat.inc"
