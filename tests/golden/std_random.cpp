// Writes known-answer vectors of the C++ standard library's std::mt19937 and
// std::uniform_real_distribution<> (libstdc++), the generators behind
// monolish's seeded vector constructor (src/utils/vector_constructor.cpp:
// 98-113).  Output: one JSON object on stdout.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

static void hexd(double v) {
  union { double d; uint64_t u; } c;
  c.d = v;
  std::printf("\"%016llx\"", (unsigned long long)c.u);
}
static void hexf(float v) {
  union { float f; uint32_t u; } c;
  c.f = v;
  std::printf("\"%08x\"", (unsigned)c.u);
}

int main() {
  std::printf("{\n");
  {
    std::mt19937 mt;  // default seed 5489
    uint32_t v = 0;
    for (int i = 0; i < 10000; i++) v = mt();
    std::printf("\"mt19937_default_10000th\": %u,\n", v);
  }
  {
    std::mt19937 mt(1234);
    std::printf("\"mt19937_1234_first\": [");
    for (int i = 0; i < 8; i++) std::printf("%s%u", i ? ", " : "", (unsigned)mt());
    std::printf("],\n");
  }
  const uint32_t seeds[] = {1001, 2001, 3001, 4000, 5001, 1234, 0, 4294967295u};
  const double ranges[][2] = {{-1.0, 1.0}, {0.0, 1.0}, {-3.5, 10.25}};
  std::printf("\"uniform_double\": [\n");
  bool first = true;
  for (uint32_t s : seeds)
    for (auto &r : ranges) {
      std::mt19937 mt(s);
      std::uniform_real_distribution<> rand(r[0], r[1]);
      std::printf("%s  {\"seed\": %u, \"min\": %.17g, \"max\": %.17g, \"values\": [",
                  first ? "" : ",\n", s, r[0], r[1]);
      first = false;
      for (int i = 0; i < 64; i++) {
        if (i) std::printf(", ");
        hexd(rand(mt));
      }
      std::printf("]}");
    }
  std::printf("\n],\n\"uniform_float\": [\n");
  first = true;
  for (uint32_t s : seeds) {
    std::mt19937 mt(s);
    std::uniform_real_distribution<> rand(-1.0f, 1.0f);
    std::printf("%s  {\"seed\": %u, \"min\": -1.0, \"max\": 1.0, \"values\": [", first ? "" : ",\n", s);
    first = false;
    for (int i = 0; i < 64; i++) {
      if (i) std::printf(", ");
      float f = rand(mt);
      hexf(f);
    }
    std::printf("]}");
  }
  std::printf("\n],\n");
  {
    // the right-hand side of the 2D Poisson 256x256 CG workload
    std::mt19937 mt(1001);
    std::uniform_real_distribution<> rand(-1.0, 1.0);
    double s = 0, s2 = 0;
    std::vector<double> v(65536);
    for (auto &x : v) x = rand(mt);
    for (double x : v) { s += x; s2 += x * x; }
    std::printf("\"rhs_65536_seed1001\": {\"last\": ");
    hexd(v.back());
    std::printf(", \"sum\": %.17g, \"sumsq\": %.17g}\n", s, s2);
  }
  std::printf("}\n");
  return 0;
}
