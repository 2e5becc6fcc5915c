// mqe_internal.h -- declarations shared by the host set-up code and the engine.
#ifndef MQE_INTERNAL_H
#define MQE_INTERNAL_H

namespace mqe {
void set_error(const char *msg);
void host_svd(int m, int n, const double *a, double *s, double *u, double *vt);
void host_qr(int m, int n, double *a, double *r);
}  // namespace mqe

#endif
