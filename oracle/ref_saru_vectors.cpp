// Test-side driver: emits Saru PRNG vectors from the reference's own
// header (include/saruprng.hpp, compiled where it lies under the reference
// tree by oracle/build_ref.sh).  Nothing in the product links this; its
// output is committed as tests/golden/saru_ref_vectors.txt and is used to pin
// the oracle's restatement of Saru (oracle/mdpd_oracle.c).
//
// Output, one record per line:
//   U <s1> <s2> <s3> <u32_0> ... <u32_7>          8 consecutive u32() draws
//   N <s1> <s2> <s3> <nrml_as_%a>                 one nrml() of a fresh stream
//   O <s1> <s2> <s3> <uni_open_0_as_%a> <uni_open_1_as_%a>
// The N record is the pair noise of nbdf_calculator.hpp:104-105, so it also
// fixes which of the two uni_open() draws feeds log and which feeds cos.
#include <cstdio>
#include <cstdint>
#include "saruprng.hpp"

static void emit(unsigned s1, unsigned s2, unsigned s3) {
  {
    Saru s(s1, s2, s3);
    std::printf("U %u %u %u", s1, s2, s3);
    for (int k = 0; k < 8; ++k) std::printf(" %u", s.u32());
    std::printf("\n");
  }
  {
    Saru s(s1, s2, s3);
    const double x = s.nrml();
    std::printf("N %u %u %u %a\n", s1, s2, s3, x);
  }
  {
    Saru s(s1, s2, s3);
    const double a = s.uni_open();
    const double b = s.uni_open();
    std::printf("O %u %u %u %a %a\n", s1, s2, s3, a, b);
  }
}

int main() {
  // edge seeds
  const unsigned edge[] = {0u, 1u, 0x7fffffffu, 0x80000000u, 0xffffffffu,
                           0xdeadbeefu};
  for (unsigned a : edge)
    for (unsigned b : edge)
      for (unsigned c : edge) emit(a, b, c);
  // the pair/step grid of test/prng_test/prng_test.cpp:17-28 at a fixed seed
  const unsigned m_seed = 100u;
  for (unsigned t = 0; t < 2; ++t)
    for (unsigned i = 0; i < 16; ++i)
      for (unsigned j = i + 1; j < 16; ++j) emit(i, j, m_seed + t);
  // a pseudo-random sweep (LCG over the seed space)
  uint32_t x = 2463534242u;
  for (int k = 0; k < 300; ++k) {
    x = x * 1664525u + 1013904223u; const unsigned a = x;
    x = x * 1664525u + 1013904223u; const unsigned b = x;
    x = x * 1664525u + 1013904223u; const unsigned c = x;
    emit(a, b, c);
  }
  return 0;
}
