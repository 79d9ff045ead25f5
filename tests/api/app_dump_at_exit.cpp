// Link-time hook for running an OP-PIC application main unchanged: with `-Wl,--wrap=opp_exit`, the
// main's call to opp_exit() lands here first. When OPPX_APP_DUMP names a directory, every particle
// dat is written there as <dat name>.bin (set->size elements of dat->size bytes each, host AoS
// order) after the set is downloaded; then the library's own opp_exit runs. It uses only the
// opp_lib API and registry (opp_dats, opp_download_particle_set), so the same file links against the
// reference's seq library and against libopp_mi355x.so, and the two dumps compare byte for byte.

#include <cstdio>
#include <cstdlib>
#include <string>

#include "opp_lib.h"

extern "C" void __real__Z8opp_exitv();

extern "C" void __wrap__Z8opp_exitv() {
    const char* dir = std::getenv("OPPX_APP_DUMP");
    if (dir && *dir) {
        for (opp_dat d : opp_dats) {
            if (!d || !d->set || !d->set->is_particle) continue;
            opp_download_particle_set(d->set, true);
            const std::string path = std::string(dir) + "/" + d->name + ".bin";
            FILE* f = std::fopen(path.c_str(), "wb");
            const size_t n = (size_t)d->set->size;
            if (!f || (n && std::fwrite(d->data, (size_t)d->size, n, f) != n)) {
                std::fprintf(stderr, "app dump: cannot write %s\n", path.c_str());
                std::exit(3);
            }
            std::fclose(f);
        }
    }
    __real__Z8opp_exitv();
}
