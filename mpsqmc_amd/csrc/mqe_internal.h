// mqe_internal.h -- declarations shared by the host set-up code and the engine.
#ifndef MQE_INTERNAL_H
#define MQE_INTERNAL_H

#include <hip/hip_runtime_api.h>

#include "mqe.h"

struct mqe_mlp;

namespace mqe {
void set_error(const char *msg);
// learned-proposal network (mqe_mlp.hip)
int mlp_forward_device(mqe_mlp *m, const float *dX, long max_rows, const int *rows_dev, int mult, float *dY,
                       hipStream_t s);
int mlp_fit_device(mqe_mlp *m, const float *dX, const float *dY, long n, int epochs, int batch, double lr,
                   int standardize, double *loss_out, hipStream_t s);
int mlp_io_dims(const mqe_mlp *m, int *nin, int *nout);
void host_svd(int m, int n, const double *a, double *s, double *u, double *vt);
void host_qr(int m, int n, double *a, double *r);
// mqe_overlaps in a caller-provided device workspace: the bytes it needs for
// all walkers at once and the least it can run with (walkers are processed in
// chunks of as many transfer-matrix workspaces as fit)
int overlaps_workspace(mqe_engine *e, const mqe_mps *other, long *bytes_all, long *bytes_min);
int overlaps_in_workspace(mqe_engine *e, const mqe_mps *other, double *out, char *ws, long ws_bytes);
}  // namespace mqe

#endif
