// lscomp_cli.h -- the command-line drivers with the reference CLIs' interface
// (lsCOMP/example/lsCOMP_xpcs.cpp:10-190 for uint16, lsCOMP_cssi.cpp:10-190
// for uint32; the two differ only in the element type and the usage example):
// the same flags (-i file -d x y z -b x y z w -p th [-o file]), the same
// validation messages and exit codes, 3 warm-up compressions, then one timed
// compression and decompression, and the same stdout lines that the AE
// scripts parse ("cuLSZ compression   end-to-end speed: %f GB/s", three
// spaces; speeds in 2^20 bytes per ms; ratio = input bytes / stream bytes).
// Timing brackets one call of the reference-named entry point with hipEvents
// on the call's stream.  -o writes the decompressed volume.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lscomp.h"
#include "lscomp_culsz.h"

#define HIP_OK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

// Element type traits: the reference-named entry points and the worst-case
// stream size of each path.
template <typename T> struct CliPath;
template <> struct CliPath<uint16_t> {
    static void compress(uint16_t *d, unsigned char *c, size_t *n, uint3 dims, uint4 bins, float p, hipStream_t s)
    { cuLSZ_compression_uint16_bsize64(d, c, n, dims, bins, p, s); }
    static void decompress(uint16_t *d, unsigned char *c, size_t n, uint3 dims, uint4 bins, float p, hipStream_t s)
    { cuLSZ_decompression_uint16_bsize64(d, c, n, dims, bins, p, s); }
    static size_t max_bytes(lscomp_dim3 d) { return lscomp_max_compressed_bytes_u16(d); }
};
template <> struct CliPath<uint32_t> {
    static void compress(uint32_t *d, unsigned char *c, size_t *n, uint3 dims, uint4 bins, float p, hipStream_t s)
    { cuLSZ_compression_uint32_bsize64(d, c, n, dims, bins, p, s); }
    static void decompress(uint32_t *d, unsigned char *c, size_t n, uint3 dims, uint4 bins, float p, hipStream_t s)
    { cuLSZ_decompression_uint32_bsize64(d, c, n, dims, bins, p, s); }
    static size_t max_bytes(lscomp_dim3 d) { return lscomp_max_compressed_bytes_u32(d); }
};

static void usage(const char *prog, const char *example)
{
    std::fprintf(stderr, "Usage:\n");
    std::fprintf(stderr, "   %s -i oriFilePath -d dims.x dims.y dims.z -b quantBins.x quantBins.y quantBins.z quantBins.w -p value -o decFilePath\n", prog);
    std::fprintf(stderr, "Options:\n");
    std::fprintf(stderr, "   -i oriFilePath: Path to the original data file\n");
    std::fprintf(stderr, "   -d dims.x dims.y dims.z: Dimensions of the original data, where dim.z is the fastest dimension.\n");
    std::fprintf(stderr, "   -b quantBins.x quantBins.y quantBins.z quantBins.w: Quantization bins for the 4 levels, where x is the base one and x<=y<=z<=w.\n");
    std::fprintf(stderr, "   -p value: Pooling threshold for a data block.\n");
    std::fprintf(stderr, "   -o decFilePath: Path to the decompressed data file (optional).\n");
    std::fprintf(stderr, "Examples:\n");
    std::fprintf(stderr, "   %s -i data/%s.bin -d %s -b 3 5 10 15 -p 0.5\n", prog, example,
                 example[0] == 'c' ? "600 1813 1558" : "1024 1813 1558");
    std::fprintf(stderr, "   %s -i data/%s.bin -d %s -b 3 5 10 15 -p 0.5 -o data/%s-dec.bin\n", prog, example,
                 example[0] == 'c' ? "600 1813 1558" : "1024 1813 1558", example);
}

// The whole CLI for element type T; `example` names the data set in the usage text.
template <typename T>
int cli_main(int argc, char *argv[], const char *example)
{
    char ori_path[640] = {0}, dec_path[640] = {0};
    uint3 dims = {0, 0, 0};
    uint4 bins = {0, 0, 0, 0};
    float pooling = 0.0f;

    if (argc < 13) {
        usage(argv[0], example);
        return 1;
    }
    for (int i = 1; i < argc; i++) {
        if (std::strcmp(argv[i], "-i") == 0) {
            if (i + 1 >= argc) { std::fprintf(stderr, "Error: Missing value for -i\n"); return 1; }
            std::strncpy(ori_path, argv[++i], sizeof(ori_path) - 1);
        } else if (std::strcmp(argv[i], "-d") == 0) {
            if (i + 3 >= argc) { std::fprintf(stderr, "Error: Missing values for -d\n"); return 1; }
            int a = std::atoi(argv[i + 1]), b = std::atoi(argv[i + 2]), c = std::atoi(argv[i + 3]);
            if (a <= 0 || b <= 0 || c <= 0) {
                std::fprintf(stderr, "Error: -d values must be positive integers\n");
                return 1;
            }
            dims.x = (unsigned)a; dims.y = (unsigned)b; dims.z = (unsigned)c;
            i += 3;
        } else if (std::strcmp(argv[i], "-b") == 0) {
            if (i + 4 >= argc) { std::fprintf(stderr, "Error: Missing values for -b\n"); return 1; }
            int q[4];
            for (int k = 0; k < 4; k++) q[k] = std::atoi(argv[i + 1 + k]);
            if (q[0] <= 0 || q[1] <= 0 || q[2] <= 0 || q[3] <= 0) {
                std::fprintf(stderr, "Error: -b values must be positive integers\n");
                return 1;
            }
            if (!(q[0] <= q[1] && q[1] <= q[2] && q[2] <= q[3])) {
                std::fprintf(stderr, "Error: -b values must satisfy x <= y <= z <= w\n");
                return 1;
            }
            bins.x = (unsigned)q[0]; bins.y = (unsigned)q[1]; bins.z = (unsigned)q[2]; bins.w = (unsigned)q[3];
            i += 4;
        } else if (std::strcmp(argv[i], "-p") == 0) {
            if (i + 1 >= argc) { std::fprintf(stderr, "Error: Missing value for -p\n"); return 1; }
            pooling = (float)std::atof(argv[++i]);
            if (pooling < 0.0f || pooling > 1.0f) {
                std::fprintf(stderr, "Error: -p value must be between 0 and 1\n");
                return 1;
            }
        } else if (std::strcmp(argv[i], "-o") == 0) {
            if (i + 1 >= argc) { std::fprintf(stderr, "Error: Missing value for -o\n"); return 1; }
            std::strncpy(dec_path, argv[++i], sizeof(dec_path) - 1);
        } else {
            std::fprintf(stderr, "Error: Unknown argument %s\n", argv[i]);
            return 1;
        }
    }

    // read the raw little-endian uint16 file
    std::vector<T> host;
    size_t n_ele = 0;
    if (FILE *f = std::fopen(ori_path, "rb")) {
        std::fseek(f, 0, SEEK_END);
        long len = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        n_ele = len > 0 ? (size_t)len / sizeof(T) : 0;
        host.resize(n_ele);
        if (n_ele && std::fread(host.data(), sizeof(T), n_ele, f) != n_ele) n_ele = 0;
        std::fclose(f);
    } else {
        std::fprintf(stderr, "Error: cannot open %s\n", ori_path);
    }
    if (n_ele != (size_t)dims.x * (size_t)dims.y * (size_t)dims.z) {
        std::fprintf(stderr, "Error: The number of elements in the original data does not match the dimensions\n");
        return 1;
    }

    const lscomp_dim3 d3 = {dims.x, dims.y, dims.z};
    const size_t cap = CliPath<T>::max_bytes(d3);
    T *d_ori = nullptr, *d_dec = nullptr;
    unsigned char *d_cmp = nullptr;
    HIP_OK(hipMalloc((void **)&d_ori, n_ele * sizeof(T)));
    HIP_OK(hipMemcpy(d_ori, host.data(), n_ele * sizeof(T), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc((void **)&d_dec, n_ele * sizeof(T)));
    HIP_OK(hipMalloc((void **)&d_cmp, cap));

    hipStream_t stream;
    HIP_OK(hipStreamCreate(&stream));
    hipEvent_t t0, t1;
    HIP_OK(hipEventCreate(&t0));
    HIP_OK(hipEventCreate(&t1));

    size_t cmp_size = 0;
    for (int i = 0; i < 3; i++)
        CliPath<T>::compress(d_ori, d_cmp, &cmp_size, dims, bins, pooling, stream);
    std::printf("GPU warmup finished!\n\n");

    HIP_OK(hipEventRecord(t0, stream));
    CliPath<T>::compress(d_ori, d_cmp, &cmp_size, dims, bins, pooling, stream);
    HIP_OK(hipEventRecord(t1, stream));
    HIP_OK(hipEventSynchronize(t1));
    float cmp_ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&cmp_ms, t0, t1));

    // stream to the host and back, so the ratio is checked against a real copy
    std::vector<unsigned char> cmp_host(cmp_size);
    HIP_OK(hipMemcpy(cmp_host.data(), d_cmp, cmp_size, hipMemcpyDeviceToHost));
    HIP_OK(hipMemset(d_cmp, 0, cap));
    HIP_OK(hipMemcpy(d_cmp, cmp_host.data(), cmp_size, hipMemcpyHostToDevice));

    HIP_OK(hipEventRecord(t0, stream));
    CliPath<T>::decompress(d_dec, d_cmp, cmp_size, dims, bins, pooling, stream);
    HIP_OK(hipEventRecord(t1, stream));
    HIP_OK(hipEventSynchronize(t1));
    float dec_ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&dec_ms, t0, t1));

    const double mib = n_ele * sizeof(T) / 1024.0 / 1024.0;
    std::printf("Dataset information:\n");
    std::printf("  - dims:   %u x %u x %u\n", dims.x, dims.y, dims.z);
    std::printf("  - length: %zu\n", n_ele);
    std::printf("  - size:   %f GB\n", n_ele * sizeof(T) / 1024.0 / 1024.0 / 1024.0);
    std::printf("Input arguments:\n");
    std::printf("  - quantBins: %u %u %u %u\n", bins.x, bins.y, bins.z, bins.w);
    std::printf("  - poolingTH: %f\n\n", pooling);
    std::printf("cuLSZ compression   end-to-end speed: %f GB/s\n", mib / cmp_ms);
    std::printf("cuLSZ decompression end-to-end speed: %f GB/s\n", mib / dec_ms);
    std::printf("cuLSZ compression ratio: %f\n", mib / (cmp_size / 1024.0 / 1024.0));

    int rc = 0;
    if (std::strlen(dec_path) > 0) {
        std::vector<T> dec(n_ele);
        HIP_OK(hipMemcpy(dec.data(), d_dec, n_ele * sizeof(T), hipMemcpyDeviceToHost));
        FILE *f = std::fopen(dec_path, "wb");
        if (!f || std::fwrite(dec.data(), sizeof(T), n_ele, f) != n_ele) {
            std::fprintf(stderr, "Error: cannot write %s\n", dec_path);
            rc = 1;
        }
        if (f) std::fclose(f);
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(d_ori);
    (void)hipFree(d_dec);
    (void)hipFree(d_cmp);
    (void)hipStreamDestroy(stream);
    return rc;
}
