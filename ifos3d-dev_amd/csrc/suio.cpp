// Seismogram files in SU format for device sections [ntr][ns+1] (the layout
// the shot context and the batched path use for observed and synthetic
// traces): the file format is csrc/suio_host.cpp's, the sections cross to
// and from the device here.
#include <hip/hip_runtime.h>

#include <vector>

#include "ifos_dev.h"
#include "suio_host.h"

using namespace ifos;

extern "C" int ifos_su_read(const char *path, int ntr, int ns, float *d_section)
{
    if (!path || ntr < 0 || ns < 1 || (ntr && !d_section)) {
        set_error("ifos_su_read: bad arguments");
        return IFOS_EINVAL;
    }
    if (ntr == 0) return IFOS_OK;
    std::vector<float> h((size_t)ntr * (ns + 1), 0.0f);
    char err[512] = "";
    if (su_read_host(path, ntr, ns, h.data(), err, sizeof(err)) != 0) {
        set_error("%s", err);
        return IFOS_EINVAL;
    }
    return hip_check(hipMemcpy(d_section, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice), "su read");
}

extern "C" int ifos_su_write(const char *path, const float *d_section, int ntr, int ns, const int *rec_xyz,
                             const int *rec_id, const float *src_xyz, float dx, float dy, float dz, float dt, int ndt)
{
    if (!path || ntr < 0 || ns < 1 || ns > 65535 || (ntr && (!d_section || !rec_xyz))) {
        set_error("ifos_su_write: bad arguments");
        return IFOS_EINVAL;
    }
    std::vector<float> sec((size_t)ntr * (ns + 1));
    if (ntr) {
        const int rc = hip_check(hipMemcpy(sec.data(), d_section, sec.size() * sizeof(float), hipMemcpyDeviceToHost),
                                 "su write");
        if (rc) return rc;
    }
    char err[512] = "";
    if (su_write_host(path, sec.data(), ntr, ns, rec_xyz, rec_id, src_xyz, dx, dy, dz, dt, ndt, err, sizeof(err))) {
        set_error("%s", err);
        return IFOS_EINVAL;
    }
    return IFOS_OK;
}
