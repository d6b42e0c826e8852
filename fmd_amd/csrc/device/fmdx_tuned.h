/* Compile-time shape of the pair kernels (k_force.hip), as chosen by
 * fmd_amd/autotune.py on one MI355X for the 4,000,000-atom EAM Cu step
 * (FCC 100^3, synthetic setfl, rc 5.5 A).  Each value can be overridden
 * with -D; the tuner builds its candidates that way.
 *
 *   FMDX_TILE64 / FMDX_TILE32   candidates per LDS tile, FP64 / FP32 kernels
 *   FMDX_TILE32_EAM             the same for the FP32 single-element EAM passes, whose tiles also
 *                               hold each candidate's list place (for the exact FP64 pairs)
 *   FMDX_TILE32_LJ / _MORSE     the same for the FP32 single-kind LJ and Morse passes
 *                               (FMDX_TILE32: the FP32 generic kernel)
 *   FMDX_EAM_NP64_0 / _1        pairs per flush iteration, FP64 force / density pass
 *   FMDX_EAM_NP32_0 / _1        the same for the FP32 (mixed) passes
 *   FMDX_WAVES_PER_CELL         wavefronts per home cell (1 or 2): one workgroup per cell either way
 *   FMDX_EAM_LDS                0, or the waves per workgroup of the FP32 single-element EAM passes
 *                               with the spline table resident in LDS (one workgroup per CU)
 *   FMDX_EAM_LDS_TILE           candidates per LDS tile in those passes (one tile per wave)
 *
 * Fixed by design, not tuned: the cell size (the binning grid, cellh >= rc,
 * whose exact atom order the kernels share).
 * Last tuning run: macro, choice, then step time per candidate:
 *   FMDX_TILE64        224  192: 3.986 ms  224: 3.899 ms  256: 4.053 ms
 *   FMDX_EAM_NP64_0      2  2: 3.893 ms  3: 4.083 ms
 *   FMDX_EAM_NP64_1      2  2: 3.897 ms  3: 3.902 ms
 *   FMDX_TILE32        384  320: 3.009 ms  384: 2.941 ms  448: 2.942 ms
 *   FMDX_EAM_NP32_0      3  2: 3.004 ms  3: 2.943 ms  4: 2.967 ms
 *   FMDX_EAM_NP32_1      4  2: 2.983 ms  4: 2.947 ms  6: 2.936 ms
 */
#ifndef FMDX_TUNED_H
#define FMDX_TUNED_H
#ifndef FMDX_TILE64
#define FMDX_TILE64 224
#endif
#ifndef FMDX_EAM_NP64_0
#define FMDX_EAM_NP64_0 2
#endif
#ifndef FMDX_EAM_NP64_1
#define FMDX_EAM_NP64_1 2
#endif
#ifndef FMDX_TILE32
#define FMDX_TILE32 384
#endif
#ifndef FMDX_TILE32_LJ
#define FMDX_TILE32_LJ 384
#endif
#ifndef FMDX_TILE32_MORSE
#define FMDX_TILE32_MORSE 384
#endif
#ifndef FMDX_TILE32_EAM
#define FMDX_TILE32_EAM 344
#endif
#ifndef FMDX_WAVES_PER_CELL
#define FMDX_WAVES_PER_CELL 1
#endif
#ifndef FMDX_EAM_LDS
#define FMDX_EAM_LDS 0
#endif
#ifndef FMDX_EAM_LDS_TILE
#define FMDX_EAM_LDS_TILE 184
#endif
#ifndef FMDX_EAM_NP32_0
#define FMDX_EAM_NP32_0 3
#endif
#ifndef FMDX_EAM_NP32_1
#define FMDX_EAM_NP32_1 4
#endif
#ifndef FMDX_SPLIT_MAX
#define FMDX_SPLIT_MAX 64
#endif
#ifndef FMDX_N3
#define FMDX_N3 0
#endif
#ifndef FMDX_TILE32_N3
#define FMDX_TILE32_N3 256
#endif
/* batch steps over the occupied cells only (lists compacted on the device after each
   re-binning) */
#ifndef FMDX_COMPACT
#define FMDX_COMPACT 1
#endif
#endif
