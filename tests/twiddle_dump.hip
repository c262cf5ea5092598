// Host program for tests/test_twiddle_tables.py: prints the compile-time
// twiddle tables of a few schedules (the kernels load the same tables).
#include <cstdio>

#include "fft.hpp"

using namespace namsa;

template <class F>
void dump(const char* name) {
  static constexpr typename F::Table tab = F::make_table();
  std::printf("# %s %d %d\n", name, F::NTW, F::T);
  for (int i = 0; i < F::NTW * F::T; ++i) std::printf("%.9e %.9e\n", tab.v[i].c, tab.v[i].s);
}

int main() {
  dump<Fft<128, 8, Radices<8, 4, 4>, false>>("128 8 8,4,4 0");
  dump<Fft<512, 8, Radices<8, 8, 8>, false>>("512 8 8,8,8 0");
  dump<Fft<512, 8, Radices<8, 8, 8>, true>>("512 8 8,8,8 1");
  dump<Fft<384, 12, Radices<3, 2, 4, 4, 4>, true>>("384 12 3,2,4,4,4 1");
  dump<Fft<768, 12, Radices<3, 4, 4, 4, 4>, false>>("768 12 3,4,4,4,4 0");
  dump<Fft<1024, 16, Radices<16, 8, 8>, true>>("1024 16 16,8,8 1");
  dump<Fft<2048, 16, Radices<16, 16, 8>, false>>("2048 16 16,16,8 0");
  dump<Fft<2048, 16, Radices<16, 16, 8>, true>>("2048 16 16,16,8 1");
  dump<Fft<768, 12, Radices<12, 4, 4, 4>, false>>("768 12 12,4,4,4 0");
  dump<Fft<384, 12, Radices<12, 4, 4, 2>, true>>("384 12 12,4,4,2 1");
  // radix 12 and 6 at a twiddled stage (compact: w, w^2, w^4 and w^8 stored)
  dump<Fft<768, 12, Radices<4, 4, 4, 12>, true>>("768 12 4,4,4,12 1");
  dump<Fft<384, 12, Radices<4, 4, 4, 6>, true>>("384 12 4,4,4,6 1");
  return 0;
}
