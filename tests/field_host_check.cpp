// Host build of the device field arithmetic (stwo-gpu_amd/csrc/field.hpp) for the CPU tests:
// reads commands from stdin, one per line, and prints the results.
//   fold  a0 a1 a2 a3  b0 b1 b2 b3  t  al0 al1 al2 al3            -> fold_mat
//   facc  a0..a3  b0..b3  t  al0..al3  acc0..acc3                   -> fold_mat_acc (alpha^2 from qsqr)
//   r64   hi lo                                                     -> reduce64(hi 2^32 + lo)
//   sub3  a b                                                       -> sub3
//   mdbl  x t                                                       -> mul_dbl(x, 2t) (= x t)
//   r62   hi lo / r63 hi lo                                         -> reduce62 / reduce63
//   qmul  a0..a3  b0..b3                                            -> qmul
#include <cstdio>
#include <cstring>

#include "field.hpp"

using namespace sh;

static qm31 rq() {
  unsigned a, b, c, d;
  if (scanf("%u %u %u %u", &a, &b, &c, &d) != 4) return qzero();
  return {{a, b}, {c, d}};
}
static void pq(qm31 v) { printf("%u %u %u %u\n", v.x.a, v.x.b, v.y.a, v.y.b); }

int main() {
  char cmd[16];
  while (scanf("%15s", cmd) == 1) {
    if (!strcmp(cmd, "fold") || !strcmp(cmd, "facc")) {
      const qm31 a = rq(), b = rq();
      unsigned t = 0;
      if (scanf("%u", &t) != 1) return 1;
      const qm31 al = rq();
      if (!strcmp(cmd, "fold")) {
        pq(fold_mat(a, b, t, qmat_of(al)));
      } else {
        const qm31 acc = rq();
        pq(fold_mat_acc(a, b, t, qmat_of(al), acc, qmat_of(qsqr(al))));
      }
    } else if (!strcmp(cmd, "r64")) {
      unsigned long long hi, lo;
      if (scanf("%llu %llu", &hi, &lo) != 2) return 1;
      printf("%u\n", reduce64((hi << 32) | lo));
    } else if (!strcmp(cmd, "mdbl")) {
      unsigned x, t;
      if (scanf("%u %u", &x, &t) != 2) return 1;
      printf("%u\n", mul_dbl(x, t + t));
    } else if (!strcmp(cmd, "r62") || !strcmp(cmd, "r63")) {
      unsigned long long hi, lo;
      if (scanf("%llu %llu", &hi, &lo) != 2) return 1;
      const unsigned long long y = (hi << 32) | lo;
      printf("%u\n", cmd[2] == '2' ? reduce62(y) : reduce63(y));
    } else if (!strcmp(cmd, "qmul")) {
      const qm31 a = rq(), b = rq();
      pq(qmul(a, b));
    } else if (!strcmp(cmd, "sub3")) {
      unsigned a, b;
      if (scanf("%u %u", &a, &b) != 2) return 1;
      printf("%u\n", sub3(a, b));
    } else {
      return 2;
    }
  }
  return 0;
}
