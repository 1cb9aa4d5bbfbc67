// suio_host.h -- the SU seismogram file format on host buffers (no HIP):
// readseis.c:27-66 and outseis.c:30-275 with SEIS_FORMAT 1.  A section is
// [ntr][ns+1] floats with sample j (1..ns) at index j.  Both return 0 or a
// negative IFOS_E* code with a message in err.
#ifndef IFOS_SUIO_HOST_H
#define IFOS_SUIO_HOST_H

#include <cstddef>

namespace ifos {
int su_read_host(const char *path, int ntr, int ns, float *section, char *err, size_t errlen);
int su_write_host(const char *path, const float *section, int ntr, int ns, const int *rec_xyz, const int *rec_id,
                  const float *src_xyz, float dx, float dy, float dz, float dt, int ndt, char *err, size_t errlen);
}  // namespace ifos

#endif
