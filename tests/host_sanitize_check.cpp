// host_sanitize_check.cpp -- drives the library's host-only units
// (nbody_ic.cpp, nbody_snapshot.cpp, nbody_metrics.cpp) on small and
// edge-case inputs, for a build with -fsanitize=address,undefined
// (tests/test_sanitize_cpu.py compiles it with g++, no GPU code involved).
// Checks statuses, not numbers: the parity tests check the numbers.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "nbody.h"
#include "nbody_host.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        failures++;
    }
}

struct Planes {
    std::vector<double> m, x, y, z, vx, vy, vz, fx, fy, fz;
    explicit Planes(size_t n)
        : m(n ? n : 1), x(n ? n : 1), y(n ? n : 1), z(n ? n : 1), vx(n ? n : 1), vy(n ? n : 1), vz(n ? n : 1),
          fx(n ? n : 1), fy(n ? n : 1), fz(n ? n : 1) {}
};

void system_calls(size_t n, const std::string& dir) {
    Planes P(n), Q(n);
    double* pos[3] = {P.x.data(), P.y.data(), P.z.data()};
    double* vel[3] = {P.vx.data(), P.vy.data(), P.vz.data()};
    double* frc[3] = {P.fx.data(), P.fy.data(), P.fz.data()};
    nbh::ic_sphere(n, 7, P.m.data(), pos, vel);
    nbh::ic_plummer(n, 7, 1.0, 20.0, P.m.data(), pos, vel);
    if (n >= 2) nbh::ic_ring(n, 1.0, 0.5, 1e5, P.m.data(), pos, vel, frc);
    const double* cpos[3] = {P.x.data(), P.y.data(), P.z.data()};
    const double* cvel[3] = {P.vx.data(), P.vy.data(), P.vz.data()};
    const double* cfrc[3] = {P.fx.data(), P.fy.data(), P.fz.data()};
    double out6[6], out2[2], I = 0.0;
    expect(nb_force_error(n, cfrc, cfrc, out6) == NB_OK || n == 0, "force_error");
    expect(nb_deviation(n, cpos, cpos, out2) == NB_OK || n == 0, "deviation");
    expect(nb_inertia(n, P.m.data(), cpos, &I) == NB_OK || n == 0, "inertia");
    if (n >= 2) expect(nb_ring_predict(n, 1.0, 0.5, 3, 100, pos, vel) == NB_OK, "ring_predict");
    // snapshots: two frames, count them, read the second back, then out-of-range and bad arguments
    const std::string path = dir + "/NBody_" + std::to_string(n) + ".dat";
    expect(nb_snapshot_write(path.c_str(), 0, 0.0, n, P.m.data(), cpos, cvel, cfrc) == NB_OK, "snapshot write");
    expect(nb_snapshot_write(path.c_str(), 1, 0.5, n, P.m.data(), cpos, cvel, nullptr) == NB_OK,
           "snapshot append");
    size_t frames = 0;
    expect(nb_snapshot_frames(path.c_str(), n, &frames) == NB_OK && frames == 2, "snapshot frames");
    double t = -1.0;
    double* qpos[3] = {Q.x.data(), Q.y.data(), Q.z.data()};
    double* qvel[3] = {Q.vx.data(), Q.vy.data(), Q.vz.data()};
    double* qfrc[3] = {Q.fx.data(), Q.fy.data(), Q.fz.data()};
    expect(nb_snapshot_read(path.c_str(), n, 1, &t, Q.m.data(), qpos, qvel, qfrc) == NB_OK && t == 0.5,
           "snapshot read");
    expect(nb_snapshot_read(path.c_str(), n, 1, nullptr, nullptr, nullptr, nullptr, nullptr) == NB_OK,
           "snapshot read, outputs skipped");
    expect(nb_snapshot_read(path.c_str(), n, 2, &t, Q.m.data(), qpos, qvel, qfrc) != NB_OK,
           "snapshot read past the last frame");
    expect(nb_snapshot_frames((dir + "/missing.dat").c_str(), n, &frames) != NB_OK, "missing file");
    std::remove(path.c_str());
}

}  // namespace

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    for (size_t n : {0, 1, 2, 3, 17, 257}) system_calls(n, dir);
    for (int order = 0; order <= 10; order++) {
        double c[16], d[16], a[49], b[8];   // the room the header documents
        (void)nbh::yoshida(order, c, d);
        (void)nbh::runge_kutta(order, a, b);
    }
    double out6[6];
    expect(nb_force_error(4, nullptr, nullptr, out6) != NB_OK, "null planes refused");
    if (failures) return 1;
    std::printf("host sanitizer run: ok\n");
    return 0;
}
