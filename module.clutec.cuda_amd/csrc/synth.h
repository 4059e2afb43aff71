/* Seeded synthetic stereo pair generator (host C).  See synth.c for the recipe. */
#ifndef CLUTEC_STEREO_SYNTH_H
#define CLUTEC_STEREO_SYNTH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SYNTH_CONSTANT = 0, SYNTH_SLANTED = 1, SYNTH_MIXED = 2 };

int stereo_synth_version(void);

/* Writes W*H bytes into L and R (row-major, pitch W) and, when gt != NULL, the
 * left-view ground-truth disparity (int16, pitch W).  Returns 0 on success. */
int stereo_synth_pair(uint64_t seed, int W, int H, int kind, int nregions, int dmax, int dbg,
                      uint8_t* L, uint8_t* R, int16_t* gt);

#ifdef __cplusplus
}
#endif
#endif
