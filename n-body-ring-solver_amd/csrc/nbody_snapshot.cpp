// nbody_snapshot.cpp -- state snapshots in the reference's binary layout.
//
// NBody_system::save_particles (CPU/nbody_system.cpp:133-152) appends one
// frame per call to NBody_<N>.dat: the time, then mass, x, y, z, vx, vy, vz,
// fx, fy, fz, N doubles each, in host byte order.  The reference has no
// reader; these routines write and read that layout so a run here can be
// checkpointed, resumed, or compared with the reference's files.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "nbody.h"
#include "nbody_host.h"

namespace {

size_t frame_doubles(size_t n) { return 1 + 10 * n; }

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

}  // namespace

extern "C" {

nb_status nb_snapshot_write(const char* path, int append, double t, size_t n, const double* mass,
                            const double* const pos[3], const double* const vel[3], const double* const force[3]) {
    if (!path || (n && (!mass || !pos || !vel || !pos[0] || !pos[1] || !pos[2] || !vel[0] || !vel[1] || !vel[2])))
        return nbh::set_error(NB_ERR_ARG, "nb_snapshot_write: null argument");
    if (force && (!force[0] || !force[1] || !force[2]))
        return nbh::set_error(NB_ERR_ARG, "nb_snapshot_write: null force plane");
    File fh;
    fh.f = fopen(path, append ? "ab" : "wb");
    if (!fh.f) return nbh::set_error(NB_ERR_ARG, std::string("nb_snapshot_write: cannot open ") + path + ": " + strerror(errno));
    bool ok = fwrite(&t, sizeof(double), 1, fh.f) == 1;
    auto put = [&](const double* p) {
        if (ok && n) ok = fwrite(p, sizeof(double), n, fh.f) == n;
    };
    put(mass);
    for (int k = 0; k < 3; k++) put(pos[k]);
    for (int k = 0; k < 3; k++) put(vel[k]);
    if (force) {
        for (int k = 0; k < 3; k++) put(force[k]);
    } else {
        const std::vector<double> zero(n, 0.0);
        for (int k = 0; k < 3; k++) put(zero.data());
    }
    ok = ok && fflush(fh.f) == 0;
    if (!ok) return nbh::set_error(NB_ERR_DEVICE, std::string("nb_snapshot_write: write failed on ") + path);
    return NB_OK;
}

nb_status nb_snapshot_frames(const char* path, size_t n, size_t* frames) {
    if (!path || !frames) return nbh::set_error(NB_ERR_ARG, "nb_snapshot_frames: null argument");
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) return nbh::set_error(NB_ERR_ARG, std::string("nb_snapshot_frames: cannot open ") + path + ": " + strerror(errno));
    if (fseek(fh.f, 0, SEEK_END) != 0) return nbh::set_error(NB_ERR_ARG, "nb_snapshot_frames: cannot seek");
    const long size = ftell(fh.f);
    const size_t fbytes = frame_doubles(n) * sizeof(double);
    if (size < 0 || size_t(size) % fbytes != 0)
        return nbh::set_error(NB_ERR_ARG, "nb_snapshot_frames: file size is not a whole number of frames for this N");
    *frames = size_t(size) / fbytes;
    return NB_OK;
}

nb_status nb_snapshot_read(const char* path, size_t n, size_t frame, double* t, double* mass, double* const pos[3],
                           double* const vel[3], double* const force[3]) {
    size_t frames = 0;
    if (nb_status s = nb_snapshot_frames(path, n, &frames)) return s;
    if (frame >= frames) return nbh::set_error(NB_ERR_ARG, "nb_snapshot_read: frame beyond the end of the file");
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) return nbh::set_error(NB_ERR_ARG, std::string("nb_snapshot_read: cannot open ") + path);
    if (fseek(fh.f, long(frame * frame_doubles(n) * sizeof(double)), SEEK_SET) != 0)
        return nbh::set_error(NB_ERR_ARG, "nb_snapshot_read: cannot seek");
    std::vector<double> buf(frame_doubles(n));
    if (fread(buf.data(), sizeof(double), buf.size(), fh.f) != buf.size())
        return nbh::set_error(NB_ERR_ARG, "nb_snapshot_read: short read");
    if (t) *t = buf[0];
    auto get = [&](int plane, double* dst) {
        if (dst && n) memcpy(dst, buf.data() + 1 + size_t(plane) * n, n * sizeof(double));
    };
    get(0, mass);
    for (int k = 0; k < 3; k++) get(1 + k, pos ? pos[k] : nullptr);
    for (int k = 0; k < 3; k++) get(4 + k, vel ? vel[k] : nullptr);
    for (int k = 0; k < 3; k++) get(7 + k, force ? force[k] : nullptr);
    return NB_OK;
}

}  // extern "C"
