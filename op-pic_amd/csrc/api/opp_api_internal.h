// Internal state shared by the opp_lib API runtime and the library's generated-loop entry points.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "opp_lib.h"
#include "oppx.h"

namespace oppx_api {

// device-side bookkeeping of one particle set
struct ParticleDev {
    oppx_move_counters* counters = nullptr;  // size / removed / hop statistics
    int size_on_device = -1;                  // set->size value last written to counters->size
    uint64_t* hole_bits = nullptr;           // removal bitmap of the last move
    void* scratch = nullptr;                 // hole-fill / sort scratch
    size_t scratch_bytes = 0;
    int64_t cap_scratch = -1;                // capacity the scratch buffers were sized for
    int32_t* hops = nullptr;
};

ParticleDev& pdev(opp_set set);
void check(int rc, const char* what);
void* dev_alloc(size_t bytes);
void dev_free(void* p);
int elem_bytes(opp_dat dat);
void sync_size_to_device(opp_set set);
void read_counters(opp_set set, oppx_move_counters* out);
void ensure_capacity_scratch(opp_set set);
void args_to_device(int nargs, opp_arg* args);
void args_written_on_device(int nargs, opp_arg* args);

// implemented with the loops
void finalize_move_device(opp_set set);
void sort_particles_device(opp_set set, bool drop_removed);
void free_loop_caches();

}  // namespace oppx_api
