// A C++ serving loop over the oppx arena (include/oppx.h), with no torch and no Python: instances
// arrive at given iterations, are admitted while they fit (first come, first served), are stepped
// together, and are retired after their own number of steps, their state written out.
//
// usage: arena_app <dir> <n_instances> <max_cells> <max_nodes> <max_particles> <max_overlay> <max_insts>
// <dir>/inst<i>/: meta.txt ("n_cells n_nodes n_part has_overlay dx dy dz minx miny minz inv_gs charge mass
// dt steps arrival"), c_det.bin c_vol.bin c2c.bin c2n.bin c_ef.bin pos.bin vel.bin p2c.bin [overlay.bin]
// writes <dir>/out<i>/{n.txt ("particles admitted_iteration retired_iteration"), pos.bin, vel.bin, lc.bin,
// p2c.bin, charge_f32.bin}

#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <deque>
#include <string>
#include <vector>

#include "oppx.h"

namespace {

template <typename T>
std::vector<T> read_bin(const std::string& path, size_t count) {
    std::vector<T> v(count);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot read %zu values from %s\n", count, path.c_str());
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

template <typename T>
void write_bin(const std::string& path, const T* data, size_t count) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(data, sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

void check(int rc, const char* what) {
    if (rc != OPPX_OK) {
        std::fprintf(stderr, "%s failed (rc=%d): %s\n", what, rc, oppx_last_error());
        std::exit(1);
    }
}

struct Instance {
    oppx_instance desc{};
    std::vector<double> det, vol, ef, pos, vel;
    std::vector<int32_t> c2c, c2n, p2c, overlay;
    int steps = 0, arrival = 0, left = 0, slot = -1, admitted_at = -1;
};

void load(const std::string& dir, Instance& I) {
    FILE* f = std::fopen((dir + "/meta.txt").c_str(), "r");
    long long nc, nn, np, has_ov, dx, dy, dz;
    double mx, my, mz, inv, q, m, dt;
    if (!f || std::fscanf(f, "%lld %lld %lld %lld %lld %lld %lld %lf %lf %lf %lf %lf %lf %lf %d %d", &nc, &nn, &np,
                          &has_ov, &dx, &dy, &dz, &mx, &my, &mz, &inv, &q, &m, &dt, &I.steps, &I.arrival) != 16) {
        std::fprintf(stderr, "bad %s/meta.txt\n", dir.c_str());
        std::exit(2);
    }
    std::fclose(f);
    I.det = read_bin<double>(dir + "/c_det.bin", 16 * nc);
    I.vol = read_bin<double>(dir + "/c_vol.bin", nc);
    I.c2c = read_bin<int32_t>(dir + "/c2c.bin", 4 * nc);
    I.c2n = read_bin<int32_t>(dir + "/c2n.bin", 4 * nc);
    I.ef = read_bin<double>(dir + "/c_ef.bin", 3 * nc);
    I.pos = read_bin<double>(dir + "/pos.bin", 3 * np);
    I.vel = read_bin<double>(dir + "/vel.bin", 3 * np);
    I.p2c = read_bin<int32_t>(dir + "/p2c.bin", np);
    oppx_instance& d = I.desc;
    d.n_cells = nc;
    d.n_nodes = nn;
    d.n_part = np;
    d.c_det = I.det.data();
    d.c_vol = I.vol.data();
    d.c2c = I.c2c.data();
    d.c2n = I.c2n.data();
    d.c_ef = I.ef.data();
    d.pos = I.pos.data();
    d.vel = I.vel.data();
    d.p2c = I.p2c.data();
    if (has_ov) {
        I.overlay = read_bin<int32_t>(dir + "/overlay.bin", (size_t)(dx * dy * dz));
        d.overlay = I.overlay.data();
    }
    d.dh_dims[0] = dx; d.dh_dims[1] = dy; d.dh_dims[2] = dz;
    d.dh_min[0] = mx; d.dh_min[1] = my; d.dh_min[2] = mz;
    d.dh_inv_gs = inv;
    d.charge = q;
    d.mass = m;
    d.dt = dt;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 8) {
        std::fprintf(stderr, "usage: %s dir n max_cells max_nodes max_particles max_overlay max_insts\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const int n = std::atoi(argv[2]);
    check(oppx_init(0), "oppx_init");
    std::vector<Instance> inst((size_t)n);
    for (int i = 0; i < n; i++) load(dir + "/inst" + std::to_string(i), inst[(size_t)i]);

    oppx_arena* arena = nullptr;
    check(oppx_arena_create(std::atoll(argv[3]), std::atoll(argv[4]), std::atoll(argv[5]), std::atoll(argv[6]),
                            std::atoi(argv[7]), nullptr, &arena),
          "oppx_arena_create");
    std::deque<int> waiting;
    for (int i = 0; i < n; i++) waiting.push_back(i);  // in arrival order (the caller sorts them)
    std::vector<int> active;
    int done = 0, faults = 0;
    for (int iter = 0; done < n; iter++) {
        while (!waiting.empty() && inst[(size_t)waiting.front()].arrival <= iter) {  // first come, first served
            Instance& I = inst[(size_t)waiting.front()];
            int32_t slot = -1;
            const int rc = oppx_arena_admit(arena, &I.desc, &slot);
            if (rc == OPPX_ERR_NOMEM) break;  // no room now: wait for a retirement
            check(rc, "oppx_arena_admit");
            I.slot = slot;
            I.left = I.steps;
            I.admitted_at = iter;
            active.push_back(waiting.front());
            waiting.pop_front();
        }
        if (active.empty()) {
            if (waiting.empty()) break;
            if (inst[(size_t)waiting.front()].arrival <= iter) {  // refused by an empty arena: never fits
                std::fprintf(stderr, "instance %d does not fit the empty arena\n", waiting.front());
                return 3;
            }
            continue;  // nothing resident: the next arrival comes at a later iteration
        }
        check(oppx_arena_step(arena), "oppx_arena_step");
        oppx_move_counters c;
        check(oppx_arena_counters(arena, &c), "oppx_arena_counters");
        faults += c.faults;
        std::vector<int> still;
        for (int i : active) {
            Instance& I = inst[(size_t)i];
            if (--I.left > 0) {
                still.push_back(i);
                continue;
            }
            const int64_t cap = I.desc.n_part;
            std::vector<double> pos(3 * cap + 3), vel(3 * cap + 3), lc(4 * cap + 4);
            std::vector<int32_t> p2c(cap + 1);
            std::vector<float> q(I.desc.n_nodes);
            int64_t k = 0;
            check(oppx_arena_retire(arena, I.slot, cap, &k, pos.data(), vel.data(), lc.data(), p2c.data(), q.data()),
                  "oppx_arena_retire");
            const std::string out = dir + "/out" + std::to_string(i);
            mkdir(out.c_str(), 0755);
            FILE* f = std::fopen((out + "/n.txt").c_str(), "w");
            std::fprintf(f, "%lld %d %d\n", (long long)k, I.admitted_at, iter);
            std::fclose(f);
            write_bin(out + "/pos.bin", pos.data(), 3 * (size_t)k);
            write_bin(out + "/vel.bin", vel.data(), 3 * (size_t)k);
            write_bin(out + "/lc.bin", lc.data(), 4 * (size_t)k);
            write_bin(out + "/p2c.bin", p2c.data(), (size_t)k);
            write_bin(out + "/charge_f32.bin", q.data(), q.size());
            done++;
        }
        active.swap(still);
    }
    check(oppx_arena_destroy(arena), "oppx_arena_destroy");
    std::printf("arena_app: %d instances served, %d faults\n", done, faults);
    return faults ? 1 : 0;
}
