// Powers of the domain generators and the omega shift, from the context's
// power tables.  References: get_powers_of_w / get_powers_of_g
// (src/ops_complex.rs:29-86, kernels native/ops_complex.cu:9-33) and
// omega_shift / omega_shift_in_place (src/ops_complex.rs:207-258,
// native/ops_complex.cu:76-87).
//
// Exponent conventions (the reference's, value for value):
//   powers of w: omega_{2^L}^(+-index); the reference forms the exponent of
//     omega_{2^24} as index << (24 - L) and masks it to 24 bits, which is
//     omega_{2^32}^(index << (32 - L)) with u32 wrap-around here.
//   powers of g: g^(+-e) with e = (index << (24 - L)) mod 2^24: the g tables
//     of the reference cover exponents below 2^24 (src/context.rs:232-249,
//     native/context.cuh:33-42), and the 24 of its OMEGA_LOG_ORDER fixes the
//     values, so L <= 24 here as there.
//   index = offset + i, or brev_L(offset + i) with bit_reverse.
// Outputs are canonical.
#include "../../include/boojum_ntt.h"
#include "ntt_device.hpp"

namespace bc {

Tables current_tables(bool &ok);  // ntt_api.hip

namespace {

constexpr uint32_t REF_G_LOG_ORDER = 24;  // reference OMEGA_LOG_ORDER (src/context.rs:12)

__device__ __forceinline__ uint32_t power_index(uint32_t raw, uint32_t log_degree, bool bit_reverse) {
  if (!bit_reverse) return raw;
  return log_degree ? __builtin_bitreverse32(raw) >> (32 - log_degree) : 0u;
}

// Linear exponents step by multiplication: a workgroup pass covers 4096
// consecutive elements as i = base + j * 256 + t (j < 16, coalesced for each
// j), so thread t starts from one table power at base + t and multiplies by
// the uniform step power (256 positions) for each j: two multiplications per
// element instead of two table reads and a multiplication.  Exponents are u32
// multiples of the table's order 2^32 (w) or stay below 2^24 (g, checked by
// the caller), so the products are exact.
constexpr uint32_t PW_NT = 256, PW_PER = 16, PW_CHUNK = PW_NT * PW_PER;

template <bool G>
__device__ __forceinline__ uint64_t table_pow(const Tables &tab, uint32_t e, bool inverse) {
  if constexpr (G) return pow_g(tab, e, inverse);
  else return pow_w(tab, inverse ? 0u - e : e);
}

// result[i] = base^(scale * (offset + i)) from the w (G false) or g table;
// for w the exponent is taken mod 2^32, for g it must stay below 2^24.
template <bool G>
__global__ __launch_bounds__(PW_NT) void powers_linear_kernel(Tables tab, uint32_t scale, uint32_t offset,
                                                              bool inverse, uint64_t *result, uint32_t count) {
  const uint64_t step = table_pow<G>(tab, scale * PW_NT, inverse);
  for (uint64_t base = (uint64_t)blockIdx.x * PW_CHUNK; base < count; base += (uint64_t)gridDim.x * PW_CHUNK) {
    const uint32_t i0 = (uint32_t)base + threadIdx.x;
    uint64_t p = table_pow<G>(tab, scale * (offset + i0), inverse);
#pragma unroll
    for (uint32_t j = 0; j < PW_PER; ++j) {
      const uint32_t i = i0 + j * PW_NT;
      if (i < count) result[i] = p;  // an ordinary store: measured 3% faster here than a non-temporal one
      p = gl::mul(p, step);
    }
  }
}

__global__ __launch_bounds__(PW_NT) void omega_shift_kernel(Tables tab, const uint64_t *src, uint32_t scale,
                                                            uint32_t offset, bool inverse, uint64_t *dst,
                                                            uint32_t count) {
  const uint64_t step = table_pow<false>(tab, scale * PW_NT, inverse);
  for (uint64_t base = (uint64_t)blockIdx.x * PW_CHUNK; base < count; base += (uint64_t)gridDim.x * PW_CHUNK) {
    const uint32_t i0 = (uint32_t)base + threadIdx.x;
    uint64_t p = table_pow<false>(tab, scale * (offset + i0), inverse);
    uint64_t x[PW_PER];
#pragma unroll
    for (uint32_t j = 0; j < PW_PER; ++j) {
      const uint32_t i = i0 + j * PW_NT;
      x[j] = i < count ? __builtin_nontemporal_load(src + i) : 0ull;  // read once: non-temporal (4% faster)
    }
#pragma unroll
    for (uint32_t j = 0; j < PW_PER; ++j) {
      const uint32_t i = i0 + j * PW_NT;
      if (i < count) __builtin_nontemporal_store(gl::mul(x[j], p), dst + i);
      p = gl::mul(p, step);
    }
  }
}

// Bit-reversed indices are not linear: one table power per element.
template <bool G>
__global__ void powers_bitrev_kernel(Tables tab, uint32_t log_degree, uint32_t offset, bool inverse,
                                     uint64_t *result, uint32_t count) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const uint32_t idx = power_index(offset + i, log_degree, true);
    const uint32_t e = G ? (idx << (REF_G_LOG_ORDER - log_degree)) & ((1u << REF_G_LOG_ORDER) - 1u)
                         : (log_degree ? idx << (32 - log_degree) : 0u);
    result[i] = table_pow<G>(tab, e, inverse);
  }
}

// g powers whose 24-bit exponent would wrap (offset + count > 2^log_degree):
// the reference's mask applied per element.
__global__ void powers_of_g_masked_kernel(Tables tab, uint32_t log_degree, uint32_t offset, bool inverse,
                                          uint64_t *result, uint32_t count) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const uint32_t e = ((offset + i) << (REF_G_LOG_ORDER - log_degree)) & ((1u << REF_G_LOG_ORDER) - 1u);
    result[i] = pow_g(tab, e, inverse);
  }
}

dim3 chunk_grid(uint32_t count) {
  const uint32_t chunks = (count + PW_CHUNK - 1) / PW_CHUNK;
  return dim3(chunks < 16384 ? (chunks ? chunks : 1) : 16384);
}

dim3 grid_for(uint32_t count) {
  const uint32_t blocks = (count + 255) / 256;
  return dim3(blocks < 65536 ? (blocks ? blocks : 1) : 65536);
}

int finish() {
  const hipError_t err = hipGetLastError();
  return err == hipSuccess ? BC_SUCCESS : (int)err;
}

}  // namespace
}  // namespace bc

extern "C" int bc_get_powers_of_w(uint32_t log_degree, uint32_t offset, bool inverse, bool bit_reverse,
                                  uint64_t *result, uint32_t count, void *stream) {
  using namespace bc;
  if (log_degree > OMEGA_LOG_ORDER || (!result && count)) return BC_ERROR_INVALID_VALUE;
  bool ok = false;
  const Tables tab = current_tables(ok);
  if (!ok) return BC_ERROR_NOT_INITIALIZED;
  if (count == 0) return BC_SUCCESS;
  hipStream_t s = (hipStream_t)stream;
  if (bit_reverse)
    powers_bitrev_kernel<false><<<grid_for(count), 256, 0, s>>>(tab, log_degree, offset, inverse, result, count);
  else  // omega_{2^32}^((offset + i) << (32 - L))
    powers_linear_kernel<false><<<chunk_grid(count), PW_NT, 0, s>>>(
        tab, log_degree ? 1u << (32 - log_degree) : 0u, offset, inverse, result, count);
  return finish();
}

extern "C" int bc_get_powers_of_g(uint32_t log_degree, uint32_t offset, bool inverse, bool bit_reverse,
                                  uint64_t *result, uint32_t count, void *stream) {
  using namespace bc;
  if (log_degree > REF_G_LOG_ORDER || (!result && count)) return BC_ERROR_INVALID_VALUE;
  bool ok = false;
  const Tables tab = current_tables(ok);
  if (!ok) return BC_ERROR_NOT_INITIALIZED;
  if (count == 0) return BC_SUCCESS;
  hipStream_t s = (hipStream_t)stream;
  if (bit_reverse)
    powers_bitrev_kernel<true><<<grid_for(count), 256, 0, s>>>(tab, log_degree, offset, inverse, result, count);
  else if ((uint64_t)offset + count <= (1ull << log_degree))  // exponents (offset + i) 2^(24 - L) < 2^24
    powers_linear_kernel<true><<<chunk_grid(count), PW_NT, 0, s>>>(tab, 1u << (REF_G_LOG_ORDER - log_degree), offset,
                                                                   inverse, result, count);
  else
    powers_of_g_masked_kernel<<<grid_for(count), 256, 0, s>>>(tab, log_degree, offset, inverse, result, count);
  return finish();
}

extern "C" int bc_omega_shift(const uint64_t *src, uint32_t log_degree, uint32_t offset, bool inverse, uint32_t shift,
                              uint64_t *dst, uint32_t count, void *stream) {
  using namespace bc;
  if (log_degree > OMEGA_LOG_ORDER || ((!src || !dst) && count)) return BC_ERROR_INVALID_VALUE;
  bool ok = false;
  const Tables tab = current_tables(ok);
  if (!ok) return BC_ERROR_NOT_INITIALIZED;
  if (count == 0) return BC_SUCCESS;
  // omega_{2^32}^(shift (offset + i) << (32 - L)), linear in i mod 2^32
  const uint32_t scale = log_degree ? shift << (32 - log_degree) : 0u;
  omega_shift_kernel<<<chunk_grid(count), PW_NT, 0, (hipStream_t)stream>>>(tab, src, scale, offset, inverse, dst,
                                                                            count);
  return finish();
}

extern "C" int bc_omega_shift_in_place(uint64_t *values, uint32_t log_degree, uint32_t offset, bool inverse,
                                       uint32_t shift, uint32_t count, void *stream) {
  return bc_omega_shift(values, log_degree, offset, inverse, shift, values, count, stream);
}
