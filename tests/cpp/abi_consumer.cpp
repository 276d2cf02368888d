// abi_consumer.cpp -- a C++ caller of include/tfs_rpy.h in the shape of the
// reference's rpy_mobility_solver usage (cilia_sim_main.cu / mobility_solver.cpp):
// pinned host arrays of the build's Real, copy_*_to_device, the evaluate_*
// sequence of evaluate_full_mobility, copy_*_to_host, wait_for_device.  The
// sequence is written once over Real through the overloads of
// include/tfs_real.hpp and run with Real = double and Real = float, as the
// reference's two builds would call it (config.hpp:36-67).
//
// Usage: abi_consumer <case.bin> <out.bin> [f32]
//   case.bin: int32 ns, nb, seg_dof, bs_accumulate; then doubles x_segs[3ns],
//             x_blobs[3nb], f_segs[6ns], f_blobs[3nb]
//   out.bin:  Real v_segs[6ns], v_blobs[3nb] (with f32: the inputs rounded to
//             float on the way in, float velocities out)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "tfs_real.hpp"

#define CHECK(call)                                                             \
    do {                                                                        \
        int rc_ = (call);                                                       \
        if (rc_ != TFS_OK) {                                                    \
            std::fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, tfs_last_error()); \
            std::exit(2);                                                       \
        }                                                                       \
    } while (0)

template <typename Real>
int run(const int hdr[4], const std::vector<double>& in, const char* out_path) {
    const int ns = hdr[0], nb = hdr[1];
    tfs_rpy_desc d = {};
    d.num_segs = ns;
    d.num_blobs = nb;
    d.seg_dof = hdr[2];
    d.precision = TFS_PRECISION_FP64;
    d.bs_mode = hdr[3] ? TFS_BS_ACCUMULATE : TFS_BS_OVERWRITE;
    d.memory = TFS_MEMORY_HOST;
    d.seg_radius = d.blob_radius = d.mu = 1.0;

    // allocate_host_memory: pinned arrays the model objects write in place
    Real *x_segs, *x_blobs, *f_segs, *f_blobs, *v_segs, *v_blobs;
    CHECK(tfs_host_alloc(sizeof(Real) * 3 * ns, (void**)&x_segs));
    CHECK(tfs_host_alloc(sizeof(Real) * 3 * nb, (void**)&x_blobs));
    CHECK(tfs_host_alloc(sizeof(Real) * 6 * ns, (void**)&f_segs));
    CHECK(tfs_host_alloc(sizeof(Real) * 3 * nb, (void**)&f_blobs));
    CHECK(tfs_host_alloc(sizeof(Real) * 6 * ns, (void**)&v_segs));
    CHECK(tfs_host_alloc(sizeof(Real) * 3 * nb, (void**)&v_blobs));
    size_t k = 0;
    for (int i = 0; i < 3 * ns; ++i) x_segs[i] = (Real)in[k++];
    for (int i = 0; i < 3 * nb; ++i) x_blobs[i] = (Real)in[k++];
    for (int i = 0; i < 6 * ns; ++i) f_segs[i] = (Real)in[k++];
    for (int i = 0; i < 3 * nb; ++i) f_blobs[i] = (Real)in[k++];

    tfs_rpy* h = nullptr;
    CHECK(tfs_rpy_create(&d, &h));                                   // allocate_device_memory
    CHECK(tfs_real::upload(h, TFS_SEG_POSITIONS, x_segs));           // read_positions_and_forces
    CHECK(tfs_real::upload(h, TFS_SEG_FORCES, f_segs));
    CHECK(tfs_real::upload(h, TFS_BLOB_POSITIONS, x_blobs));
    CHECK(tfs_real::upload(h, TFS_BLOB_FORCES, f_blobs));
    // evaluate_full_mobility (rpy_mobility_solver.cu:330-337)
    CHECK(tfs_rpy_evaluate(h, TFS_BLOCK_SS));
    CHECK(tfs_rpy_evaluate(h, TFS_BLOCK_SB));
    CHECK(tfs_real::download(h, TFS_SEG_VELOCITIES, v_segs));
    CHECK(tfs_rpy_evaluate(h, TFS_BLOCK_BB));
    CHECK(tfs_rpy_evaluate(h, TFS_BLOCK_BS));
    CHECK(tfs_real::download(h, TFS_BLOB_VELOCITIES, v_blobs));
    CHECK(tfs_rpy_sync(h));                                          // wait_for_device

    FILE* o = std::fopen(out_path, "wb");
    if (!o) return 1;
    std::fwrite(v_segs, sizeof(Real), 6 * ns, o);
    std::fwrite(v_blobs, sizeof(Real), 3 * nb, o);
    std::fclose(o);
    CHECK(tfs_rpy_destroy(h));
    for (Real* p : {x_segs, x_blobs, f_segs, f_blobs, v_segs, v_blobs}) CHECK(tfs_host_free(p));
    std::printf("abi_consumer: %s %d segments, %d blobs, Real of %zu bytes\n", tfs_version(), ns, nb, sizeof(Real));
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 3 && argc != 4) {
        std::fprintf(stderr, "usage: %s case.bin out.bin [f32]\n", argv[0]);
        return 1;
    }
    const bool f32 = argc == 4 && std::strcmp(argv[3], "f32") == 0;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 1;
    int hdr[4];
    if (std::fread(hdr, sizeof(int), 4, f) != 4) return 1;
    const size_t n = 9 * (size_t)hdr[0] + 6 * (size_t)hdr[1];
    std::vector<double> in(n);
    const size_t got = std::fread(in.data(), sizeof(double), n, f);
    std::fclose(f);
    if (got != n) return 1;
    return f32 ? run<float>(hdr, in, argv[2]) : run<double>(hdr, in, argv[2]);
}
