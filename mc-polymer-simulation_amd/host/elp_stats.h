/*
 * elp_stats.h -- polymer statistics of ELP configurations (libelp_stats.so).
 *
 * A restatement of the reference analysis (analysis/mode_io.c:510-600 reading
 * and unwrapping, analysis/mode_compute.c:145-217 SL ratio / CMS / Rg^2,
 * :429-504 contact probability, :506-570 monomer and CMS mean-square
 * displacements, analysis/mode_io.c:54-70 normalisation), for one system:
 * n_chain chains of n_mono monomers observed over a series of frames.
 *
 * Frames come from tuv files (the format of gpupol2/gpupol_io.c:105-203 and
 * cpupol/efpol_io.c:3-77) or from arrays (start site + bond codes per chain).
 * Positions are unwrapped along each chain from its start site, the start is
 * unwrapped against the previous frame by the minimum image, and the chain
 * is then moved by whole boxes so that its mean position moves by less than
 * half a box between frames (analysis/mode_io.c:531-581).
 * Cartesian coordinates x = t+u-v, y = t-u, z = v; |bond|^2 = 2, so lengths
 * are reported in bond units (divided by sqrt 2) as the reference does.
 */
#ifndef ELP_STATS_H
#define ELP_STATS_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct elp_series elp_series;

/* ring: 1 for rings (n_mono bonds close the chain), 0 for linear chains
   (n_mono-1 bonds, the tuv text ends with 'f'). */
elp_series* elp_series_new(int ring, int LT, int LU, int LV, int n_chain, int n_mono, int max_frames);
void elp_series_free(elp_series* s);

/* Add one frame: chain c starts at starts[3c..3c+2] (t,u,v) and has bond codes
   bonds[c*stride + i], i < n_mono (0 = stored length; linear: the last entry
   is ignored).  Returns 0, or -1 when the series is full / sizes disagree. */
int elp_series_add(elp_series* s, const int32_t* starts, const uint8_t* bonds, int stride);

/* Add one frame read from a tuv file; every chain must have n_mono monomers.
   Returns 0 or a negative code (-1 full, -2 unreadable, -3 shape mismatch). */
int elp_series_add_tuv(elp_series* s, const char* path);

int elp_series_frames(const elp_series* s);

/* Unwrapped (t,u,v) of chain c in frame f: out[3*i + 0..2], i < n_mono. */
int elp_series_positions(const elp_series* s, int f, int c, int32_t* out);

/* Squared radius of gyration of every chain in frame f (bond units). */
int elp_series_rg2(const elp_series* s, int f, double* rg2);

/*
 * Equilibrium statics over frames n_therm.. and all chains:
 *   rg2    mean squared radius of gyration (ComputeGyration)
 *   slrat  mean fraction of bonds with both ends on one site (ComputeSL)
 *   pc     pc[g], g < n_mono: contact probability of monomers g apart along
 *          the chain (same or nearest-neighbour site, same periodic image),
 *          normalised by n_chain * frames * (n_mono - g) (WriteAvgContactProbability)
 */
int elp_series_statics(const elp_series* s, int n_therm, double* rg2, double* slrat, double* pc);

/*
 * Mean-square displacements over frames n_therm.. at lag k frames, k < n_out:
 *   g1[k]  middle monomer (index polSize/2), minus the drift of the whole
 *          system's mean position, |dr|^2 / 2 (AddMono)
 *   g3[k]  chain centre of mass, minus the same drift (AddCMS)
 * averaged over all time origins and chains.
 */
int elp_series_msd(const elp_series* s, int n_therm, int n_out, double* g1, double* g3);

/*
 * The same displacements with the reference analysis's drift convention
 * (analysis/mode_compute.c:244-251,506-570, analysis/mode_io.c:640-652): its
 * avg_pos file holds the mean over chains of cms/sqrt(2), which g3 subtracts
 * from a centre-of-mass displacement in bond units and g1 from a
 * middle-monomer displacement in lattice units before dividing by 2, so g3
 * removes 1/sqrt(2) of the system mean's displacement and g1 half of it.
 * Reproduces the reference's g1/g3 files from the same frames.
 */
int elp_series_msd_reference(const elp_series* s, int n_therm, int n_out, double* g1, double* g3);

/*
 * Summary mode: a series fed with per-chain reductions of each frame instead
 * of all monomer positions (what elp_ring_frame_summary of libelp.so returns):
 * per chain c its start site start3[3c..] in the box, the sum of its monomer
 * positions relative to the start sum3[3c..], the position of monomer
 * polSize/2 relative to the start mid3[3c..], rg2[c], the number of
 * stored-length bonds nsl[c]; per frame contacts[g], g < n_mono (pairs g
 * apart that touch, summed over chains).  Frames are unwrapped as in the full
 * mode, and elp_series_statics / elp_series_msd / elp_series_rg2 give the same
 * results (rg2 up to rounding: it is taken as given); positions are not kept.
 */
elp_series* elp_series_new_summary(int ring, int LT, int LU, int LV, int n_chain, int n_mono, int max_frames);
int elp_series_add_summary(elp_series* s, const int32_t* start3, const int64_t* sum3, const int32_t* mid3,
                           const double* rg2, const int32_t* nsl, const uint64_t* contacts);
/* The summary of frame f of a full-mode series (start3 wrapped into the box). */
int elp_series_summary(const elp_series* s, int f, int32_t* start3, int64_t* sum3, int32_t* mid3, double* rg2,
                       int32_t* nsl, uint64_t* contacts);

#ifdef __cplusplus
}
#endif
#endif
