// culsz_compat.cpp -- reference-named C++ entry points over the C ABI.
// Mirrors lsCOMP/src/cuLSZ_entry.cu:18-166 (names, parameter order, the
// print-and-exit error convention of cuLSZ_entry.cu:9-14).
#include <cstdio>
#include <cstdlib>

#include "../../include/lscomp.h"
#include "../../include/lscomp_culsz.h"

static void culsz_check(int st, const char *what)
{
    if (st == LSCOMP_OK) return;
    std::fprintf(stderr, "lsCOMP error in %s: code=%d \"%s\"\n", what, st, lscomp_status_string(st));
    std::exit(st);
}

void cuLSZ_compression_uint16_bsize64(uint16_t *d_oriData, unsigned char *d_cmpBytes, size_t *cmpSize,
                                      uint3 dims, uint4 quantBins, float poolingSH, hipStream_t stream)
{
    const lscomp_dim3 d = {dims.x, dims.y, dims.z};
    const lscomp_bins4 b = {quantBins.x, quantBins.y, quantBins.z, quantBins.w};
    size_t n = 0;
    int st = lscomp_compress_u16(d_oriData, d_cmpBytes, lscomp_max_compressed_bytes_u16(d), &n, d, b,
                                 poolingSH, stream);
    culsz_check(st, "cuLSZ_compression_uint16_bsize64");
    *cmpSize = n;
}

void cuLSZ_decompression_uint16_bsize64(uint16_t *d_decData, unsigned char *d_cmpBytes, size_t cmpSize,
                                        uint3 dims, uint4 quantBins, float poolingSH, hipStream_t stream)
{
    (void)poolingSH;   // unused on decode, as in the reference
    const lscomp_dim3 d = {dims.x, dims.y, dims.z};
    const lscomp_bins4 b = {quantBins.x, quantBins.y, quantBins.z, quantBins.w};
    int st = lscomp_decompress_u16(d_decData, d_cmpBytes, cmpSize, d, b, stream);
    culsz_check(st, "cuLSZ_decompression_uint16_bsize64");
}

void cuLSZ_compression_uint32_bsize64(uint32_t *d_oriData, unsigned char *d_cmpBytes, size_t *cmpSize,
                                      uint3 dims, uint4 quantBins, float poolingTH, hipStream_t stream)
{
    const lscomp_dim3 d = {dims.x, dims.y, dims.z};
    const lscomp_bins4 b = {quantBins.x, quantBins.y, quantBins.z, quantBins.w};
    size_t n = 0;
    int st = lscomp_compress_u32(d_oriData, d_cmpBytes, lscomp_max_compressed_bytes_u32(d), &n, d, b,
                                 poolingTH, stream);
    culsz_check(st, "cuLSZ_compression_uint32_bsize64");
    *cmpSize = n;
}

void cuLSZ_decompression_uint32_bsize64(uint32_t *d_decData, unsigned char *d_cmpBytes, size_t cmpSize,
                                        uint3 dims, uint4 quantBins, float poolingTH, hipStream_t stream)
{
    (void)poolingTH;   // unused on decode, as in the reference
    const lscomp_dim3 d = {dims.x, dims.y, dims.z};
    const lscomp_bins4 b = {quantBins.x, quantBins.y, quantBins.z, quantBins.w};
    int st = lscomp_decompress_u32(d_decData, d_cmpBytes, cmpSize, d, b, stream);
    culsz_check(st, "cuLSZ_decompression_uint32_bsize64");
}
