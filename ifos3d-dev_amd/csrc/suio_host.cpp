// Seismogram files in SU format (readseis.c, outseis.c with SEIS_FORMAT 1):
// each trace is a 240-byte SEG-Y/SU trace header followed by ns native
// 4-byte floats.  Host buffers only; csrc/suio.cpp moves the sections to and
// from the device.
#include "suio_host.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

namespace {

// IFOS_EINVAL of include/ifos3d_hip.h
constexpr int kEinval = -1;

// byte offsets of the header words the writer sets (the standard SU trace
// header; segy.h of the reference declares the same words in this order)
enum : int {
    H_TRACL = 0, H_TRACR = 4, H_FLDR = 8, H_TRACF = 12, H_EP = 16, H_CDPT = 24, H_TRID = 28, H_NVS = 30,
    H_NHS = 32, H_DUSE = 34, H_OFFSET = 36, H_GELEV = 40, H_SDEPTH = 48, H_GDEL = 52, H_SDEL = 56, H_GWDEP = 64,
    H_SCALEL = 68, H_SCALCO = 70, H_SX = 72, H_SY = 76, H_GX = 80, H_GY = 84, H_NS = 114, H_DT = 116,
    H_D1 = 180, H_F1 = 184, H_D2 = 188, H_F2 = 192, H_UNGPOW = 196, H_UNSCALE = 200, H_NTR = 204, H_MARK = 208,
    H_BYTES = 240
};

template <typename T>
void put(unsigned char *h, int off, T v)
{
    std::memcpy(h + off, &v, sizeof(T));
}

// iround of fd.h:35, floor(x + 0.5) in double
int iround(double x)
{
    return (int)std::floor(x + 0.5);
}

int fail(char *err, size_t errlen, const char *what, const char *path)
{
    if (err && errlen) std::snprintf(err, errlen, "%s%s", what, path ? path : "");
    return kEinval;
}

}  // namespace

namespace ifos {

int su_read_host(const char *path, int ntr, int ns, float *section, char *err, size_t errlen)
{
    if (!path || ntr < 0 || ns < 1 || (ntr && !section)) return fail(err, errlen, "ifos_su_read: bad arguments", nullptr);
    if (ntr == 0) return 0;  // readseis.c:53: nothing to read
    std::memset(section, 0, (size_t)ntr * (ns + 1) * sizeof(float));
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(err, errlen, "ifos_su_read: cannot open ", path);
    // readseis.c:56-63: skip each trace header, read ns floats; samples past
    // the end of a short file stay 0, as there (section[i][j] = 0 before
    // each fread)
    for (int i = 0; i < ntr; ++i) {
        float *row = &section[(size_t)i * (ns + 1) + 1];
        if (std::fseek(fp, H_BYTES, SEEK_CUR) != 0) break;
        const size_t got = std::fread(row, sizeof(float), (size_t)ns, fp);
        if (got < (size_t)ns) {
            std::memset(row + got, 0, ((size_t)ns - got) * sizeof(float));
            break;
        }
    }
    std::fclose(fp);
    return 0;
}

int su_write_host(const char *path, const float *section, int ntr, int ns, const int *rec_xyz, const int *rec_id,
                  const float *src_xyz, float dx, float dy, float dz, float dt, int ndt, char *err, size_t errlen)
{
    if (!path || ntr < 0 || ns < 1 || ns > 65535 || (ntr && (!section || !rec_xyz)))
        return fail(err, errlen, "ifos_su_write: bad arguments", nullptr);
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(err, errlen, "ifos_su_write: cannot open ", path);
    // outseis.c:46-56: source coordinates in the header only for one source
    const float XS = src_xyz ? src_xyz[0] : 0.0f, YS = src_xyz ? src_xyz[1] : 0.0f,
                ZS = src_xyz ? src_xyz[2] : 0.0f;
    const float scale = 3.0f;
    const float scalefac = (float)std::pow(10.0, (double)scale);
    unsigned char h[H_BYTES];
    int rc = 0;
    for (int t = 0; t < ntr && !rc; ++t) {
        std::memset(h, 0, sizeof(h));  // words outseis.c leaves unset are written as 0
        // outseis.c:62-101 (the words it later resets to 0 stay 0)
        const float xr = rec_xyz[3 * t] * dx, yr = rec_xyz[3 * t + 1] * dy, zr = rec_xyz[3 * t + 2] * dz;
        put<int32_t>(h, H_TRACL, rec_id ? rec_id[t] : t + 1);
        put<int16_t>(h, H_TRID, 1);
        put<int32_t>(h, H_OFFSET,
                     iround(std::sqrt((double)((XS - xr) * (XS - xr) + (YS - yr) * (YS - yr) + (ZS - zr) * (ZS - zr))) *
                            scalefac));
        put<int32_t>(h, H_GELEV, iround(yr * scalefac));
        put<int32_t>(h, H_SDEPTH, iround(YS * scalefac));
        put<int16_t>(h, H_SCALEL, (int16_t)(-scale));
        put<int16_t>(h, H_SCALCO, (int16_t)(-scale));
        put<int32_t>(h, H_SX, iround(XS * scalefac));
        put<int32_t>(h, H_SY, iround(ZS * scalefac));
        put<int32_t>(h, H_GX, iround(xr * scalefac));
        put<int32_t>(h, H_GY, iround(zr * scalefac));
        put<uint16_t>(h, H_NS, (uint16_t)ns);
        put<uint16_t>(h, H_DT, (uint16_t)iround(((float)ndt * dt) * 1.0e6));
        if (std::fwrite(h, 1, H_BYTES, fp) != H_BYTES ||
            std::fwrite(&section[(size_t)t * (ns + 1) + 1], sizeof(float), (size_t)ns, fp) != (size_t)ns)
            rc = fail(err, errlen, "ifos_su_write: write failed: ", path);
    }
    std::fclose(fp);
    return rc;
}

}  // namespace ifos
