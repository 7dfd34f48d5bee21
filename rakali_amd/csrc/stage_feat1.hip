// Stage kernels with surface forcing (FEAT = 1).
#define RKG_FEAT 1
#include "stage_feat.inc"
