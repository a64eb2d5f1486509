/*
 * dpk_deeppot.h -- C API of libdpk_deeppot, the C++ MD-engine runner of a
 * se_e2_a model on libdpk (MI355X): the counterpart of deepmd::DeepPotPT
 * (source/api_cc/src/DeepPotPT.cc) for an engine that hands host arrays and
 * its own neighbour list.  It drives the C ABI of dpk.h only: list
 * formatting, environment matrix, embedding, fused split-FP16 fitting,
 * embedding backward, forces / virial, frame sums.
 *
 * Models come from deepmd_amd.export.export_lower_model (split-FP16 models
 * with the fused fitting).  Host pointers everywhere; all work runs on a
 * stream owned by the handle, and dpk_deeppot_compute returns when the
 * results are on the host.  One handle per host thread at a time.
 */
#ifndef DPK_DEEPPOT_H
#define DPK_DEEPPOT_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dpk_deeppot dpk_deeppot;

/* Replaces DeepPotPT::init (DeepPotPT.cc:47-110): reads the exported model
 * and uploads it to GPU `device`.  NULL on failure (see
 * dpk_deeppot_last_error). */
dpk_deeppot* dpk_deeppot_load(const char* path, int device);
void dpk_deeppot_free(dpk_deeppot* dp);

/* Replaces DeepPotPT::compute with an InputNlist (DeepPotPT.cc:171-378):
 *   coord [nall * 3]: local atoms first, then the nghost ghost atoms;
 *   atype [nall]: types in [0, ntypes) (the engine removes NULL-type atoms);
 *   the engine's list in CSR form (InputNlist: inum rows, row ii is local
 *   atom ilist[ii] with numneigh[ii] neighbours firstneigh[ii][...], indices
 *   into coord; any order, any cutoff >= rcut), read when ago == 0 and kept
 *   while ago > 0 (between re-neighbourings the local atoms, their types and
 *   the list stay; coordinates change);
 *   energy [1], force [nall * 3] (extended: ghost forces are for the engine
 *   to add to their owners), virial [9].
 * Returns 0 or an error code of dpk.h. */
int dpk_deeppot_compute(dpk_deeppot* dp, double* energy, double* force, double* virial,
                        const double* coord, const int* atype, int nall, int nghost, int inum,
                        const int* ilist, const int* numneigh, const int* const* firstneigh,
                        int ago);

double dpk_deeppot_rcut(const dpk_deeppot* dp);
int dpk_deeppot_ntypes(const dpk_deeppot* dp);
const char* dpk_deeppot_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
