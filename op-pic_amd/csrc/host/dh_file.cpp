// Direct-hop overlay files: the reference's on-disk format, so overlays move between this build and
// OP-PIC in both directions. The file is the zlib stream (compress(), default level) of the
// overlay's int32 entries in native byte order (opp_lib/src/core/opp_util.cpp:442-496), named
//   dh_c<cells>_gs<spacing %2.4lE>_ex<expansion.x %2.4lE>_<expansion.y>_<expansion.z>.bin
// (opp_lib/src/seq/opp_direct_hop_seq.cpp:107-119 read, :288-298 write).

#include <zlib.h>

#include <cstdio>
#include <fstream>
#include <vector>

#include "oppx.h"

extern "C" {

int oppx_dh_overlay_filename(int64_t n_cells, double grid_spacing, const double expansion[3], char* out,
                             size_t cap) {
    if (!expansion || !out || cap == 0 || n_cells < 0) return OPPX_ERR_ARG;
    const int w = std::snprintf(out, cap, "dh_c%d_gs%2.4lE_ex%2.4lE_%2.4lE_%2.4lE.bin", (int)n_cells, grid_spacing,
                                expansion[0], expansion[1], expansion[2]);
    return (w > 0 && (size_t)w < cap) ? OPPX_OK : OPPX_ERR_ARG;
}

int oppx_dh_overlay_save(const char* path, const int32_t* overlay, int64_t n) {
    if (!path || !overlay || n <= 0) return OPPX_ERR_ARG;
    const uLong src = (uLong)n * sizeof(int32_t);
    uLong dst = compressBound(src);
    std::vector<Bytef> buf(dst);
    if (compress(buf.data(), &dst, reinterpret_cast<const Bytef*>(overlay), src) != Z_OK) return OPPX_ERR_STATE;
    std::ofstream f(path, std::ios::binary);
    if (!f) return OPPX_ERR_ARG;
    f.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)dst);
    return f ? OPPX_OK : OPPX_ERR_STATE;
}

int oppx_dh_overlay_load(const char* path, int32_t* overlay, int64_t n) {
    if (!path || !overlay || n <= 0) return OPPX_ERR_ARG;
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return OPPX_ERR_ARG;
    const std::streamsize sz = f.tellg();
    f.seekg(0);
    std::vector<Bytef> buf((size_t)sz);
    if (!f.read(reinterpret_cast<char*>(buf.data()), sz)) return OPPX_ERR_STATE;
    uLong dst = (uLong)n * sizeof(int32_t);
    if (uncompress(reinterpret_cast<Bytef*>(overlay), &dst, buf.data(), (uLong)sz) != Z_OK) return OPPX_ERR_STATE;
    return dst == (uLong)n * sizeof(int32_t) ? OPPX_OK : OPPX_ERR_STATE;  // the grid must match the file
}

}  // extern "C"
