/*
 * lscomp_culsz.h -- the reference's host entry points (uint16 and uint32) under
 * their own names, parameter order and C++ linkage, implemented on the gfx950
 * codec.
 *
 * A caller of lsCOMP/include/cuLSZ_entry.h:8-11 switches by including this
 * header instead and linking liblscomp.so; cudaStream_t becomes hipStream_t
 * and uint3/uint4 are HIP's own vector types.
 *
 * Behaviour kept from the reference (lsCOMP/src/cuLSZ_entry.cu:93-166): the
 * calls return when the stream size is known; a HIP or argument error prints
 * a message and exits the process with the error code (cuLSZ_entry.cu:9-14).
 * Differences (defined where the reference is undefined or wrong):
 *   - *cmpSize is the full stream size R + sum(payload) (the reference stores
 *     sum(payload) + blocks, cuLSZ_entry.cu:125, and only blocks when the
 *     volume has at most 1024 blocks);
 *   - d_cmpBytes must hold lscomp_max_compressed_bytes_u16(dims) bytes, or at
 *     least the stream (nothing past that bound is written);
 *   - decompression writes zeros for rate-0 blocks and checks the stream
 *     against cmpSize (the reference ignores cmpSize);
 *   - uint32: a volume needing rate 32 in some block is an error
 *     (LSCOMP_ERR_RANGE; the reference writes an undecodable stream,
 *     cuLSZ_kernel.cu:120-131, 259-261), and offsets are 64-bit (the
 *     reference's are 32-bit, cuLSZ_entry.cu:31-38).
 */
#ifndef LSCOMP_CULSZ_H
#define LSCOMP_CULSZ_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime.h>

/* cuLSZ_entry.h:8 */
void cuLSZ_compression_uint32_bsize64(uint32_t *d_oriData, unsigned char *d_cmpBytes, size_t *cmpSize,
                                      uint3 dims, uint4 quantBins, float poolingTH, hipStream_t stream = 0);
/* cuLSZ_entry.h:9 */
void cuLSZ_decompression_uint32_bsize64(uint32_t *d_decData, unsigned char *d_cmpBytes, size_t cmpSize,
                                        uint3 dims, uint4 quantBins, float poolingTH, hipStream_t stream = 0);
/* cuLSZ_entry.h:10 */
void cuLSZ_compression_uint16_bsize64(uint16_t *d_oriData, unsigned char *d_cmpBytes, size_t *cmpSize,
                                      uint3 dims, uint4 quantBins, float poolingSH, hipStream_t stream = 0);
/* cuLSZ_entry.h:11 */
void cuLSZ_decompression_uint16_bsize64(uint16_t *d_decData, unsigned char *d_cmpBytes, size_t cmpSize,
                                        uint3 dims, uint4 quantBins, float poolingSH, hipStream_t stream = 0);

#endif /* LSCOMP_CULSZ_H */
