// CPU emulation of hipcub::DeviceRadixSort::SortPairs and DeviceScan (test
// infrastructure). The sort:
// a stable sort on the keys' radix order (IEEE floats ordered as the radix
// sort orders them: sign-flipped bit patterns).
#pragma once

#include <algorithm>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

#include <hip/hip_runtime.h>

namespace hipcub {

namespace detail {
template <class K>
inline unsigned long long radix_bits(K k) {
  if constexpr (std::is_floating_point<K>::value) {
    static_assert(sizeof(K) == 8, "double keys only");
    unsigned long long u;
    std::memcpy(&u, &k, 8);
    return (u >> 63) ? ~u : (u | (1ull << 63));
  } else {
    return (unsigned long long)k;
  }
}
}  // namespace detail

struct DeviceRadixSort {
  template <class K, class V>
  static hipError_t SortPairs(void* temp, size_t& temp_bytes, const K* keys_in, K* keys_out, const V* vals_in,
                              V* vals_out, int n, int begin_bit = 0, int end_bit = (int)(sizeof(K) * 8),
                              hipStream_t stream = nullptr) {
    (void)begin_bit;
    (void)end_bit;
    (void)stream;
    if (!temp) {
      temp_bytes = 16;
      return hipSuccess;
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
      return detail::radix_bits(keys_in[a]) < detail::radix_bits(keys_in[b]);
    });
    std::vector<K> k2(n);
    std::vector<V> v2(n);
    for (int i = 0; i < n; i++) {
      k2[i] = keys_in[idx[i]];
      v2[i] = vals_in[idx[i]];
    }
    std::copy(k2.begin(), k2.end(), keys_out);
    std::copy(v2.begin(), v2.end(), vals_out);
    return hipSuccess;
  }
};

// Prefix sums (inclusive and exclusive) over n items, sequentially.
struct DeviceScan {
  template <class T>
  static hipError_t InclusiveSum(void* temp, size_t& temp_bytes, const T* in, T* out, int n,
                                 hipStream_t stream = nullptr) {
    (void)stream;
    if (!temp) {
      temp_bytes = 16;
      return hipSuccess;
    }
    T acc = T(0);
    for (int i = 0; i < n; i++) {
      acc = acc + in[i];
      out[i] = acc;
    }
    return hipSuccess;
  }
  template <class T>
  static hipError_t ExclusiveSum(void* temp, size_t& temp_bytes, const T* in, T* out, int n,
                                 hipStream_t stream = nullptr) {
    (void)stream;
    if (!temp) {
      temp_bytes = 16;
      return hipSuccess;
    }
    T acc = T(0);
    for (int i = 0; i < n; i++) {
      const T v = in[i];
      out[i] = acc;
      acc = acc + v;
    }
    return hipSuccess;
  }
};

}  // namespace hipcub
