// Host AddressSanitizer / UndefinedBehaviorSanitizer driver (test
// infrastructure): the CPU build of the DCS physics (noa_amd/csrc/dcs_host.cc)
// and the oracle (oracle/dcs_oracle.cc), compiled with -fsanitize by
// tests/test_host_sanitizers.py and run on special and ordinary inputs. It
// checks: the error returns of noa_dcs_eval_f64_host; that one thread and an
// OpenMP team give the same bits (disjoint elementwise writes); that the host
// build agrees with the oracle to 1e-10 relative with the same zero and NaN
// pattern. Any sanitizer report aborts the process (-fno-sanitize-recover).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "noa_dcs.h"

extern "C" int oracle_dcs_map_all(const double* K, const double* q, int64_t n, double A, double I, int32_t Z,
                                  double mass, double* out, int threads);

static int failures = 0;

static void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        failures++;
    }
}

int main() {
    const double nan = std::numeric_limits<double>::quiet_NaN(), inf = std::numeric_limits<double>::infinity();
    // an ordinary log-uniform grid and a grid of special and extreme values
    std::vector<double> K, q;
    for (int i = 0; i < 64; i++)
        for (int j = 0; j < 64; j++) {
            const double k = std::pow(10., -3. + 9. * i / 63.);
            K.push_back(k);
            q.push_back(k * std::pow(10., -6. + 6. * j / 64.));
        }
    const double special[] = {nan, inf, -inf, 0.0, -1.0, 1e-300, 5e-324, 1e300, 1e30, 1e13, 1e-3, 0.5, 3.0, 1e6};
    for (double a : special)
        for (double b : special) {
            K.push_back(a);
            q.push_back(b);
        }
    const int64_t n = int64_t(K.size());
    const noa_dcs_element elements[] = {{1.00794, 19.2e-9, 1}, {22.0, 0.1364e-6, 11}, {238.02891, 890e-9, 92}};
    const double masses[] = {0.10565839, 1.77682};

    // error returns
    std::vector<double> o1(4 * n), o2(4 * n), ref(4 * n);
    noa_dcs_outputs_f64 out{o1.data(), o1.data() + n, o1.data() + 2 * n, o1.data() + 3 * n};
    expect(noa_dcs_eval_f64_host(K.data(), q.data(), n, elements, 0.105, nullptr, 1) != 0, "NULL outputs");
    expect(noa_dcs_eval_f64_host(K.data(), q.data(), -1, elements, 0.105, &out, 1) != 0, "n < 0");
    expect(noa_dcs_eval_f64_host(K.data(), q.data(), n, nullptr, 0.105, &out, 1) != 0, "NULL element");
    const noa_dcs_element bad{-1.0, 1e-7, 11};
    expect(noa_dcs_eval_f64_host(K.data(), q.data(), n, &bad, 0.105, &out, 1) != 0, "A < 0");
    expect(noa_dcs_eval_f64_host(K.data(), q.data(), n, elements, -1.0, &out, 1) != 0, "mass < 0");
    expect(noa_dcs_eval_f64_host(K.data(), q.data(), n, elements, 0.105, &out, -2) != 0, "threads < 0");

    for (const auto& el : elements)
        for (double mass : masses) {
            noa_dcs_outputs_f64 a{o1.data(), o1.data() + n, o1.data() + 2 * n, o1.data() + 3 * n};
            noa_dcs_outputs_f64 b{o2.data(), o2.data() + n, o2.data() + 2 * n, o2.data() + 3 * n};
            expect(noa_dcs_eval_f64_host(K.data(), q.data(), n, &el, mass, &a, 1) == 0, "serial call");
            expect(noa_dcs_eval_f64_host(K.data(), q.data(), n, &el, mass, &b, 4) == 0, "OpenMP call");
            expect(std::memcmp(o1.data(), o2.data(), sizeof(double) * o1.size()) == 0, "serial == OpenMP bits");
            expect(oracle_dcs_map_all(K.data(), q.data(), n, el.A, el.I, el.Z, mass, ref.data(), 1) == 0,
                   "oracle call");
            double worst = 0;
            for (int64_t i = 0; i < 4 * n; i++) {
                const double g = o1[i], r = ref[i];
                if (std::isnan(r) || std::isnan(g)) {
                    expect(std::isnan(r) && std::isnan(g), "NaN pattern");
                } else if (r == 0 || g == 0 || std::isinf(r)) {
                    expect(g == r, "zero / infinity pattern");
                } else {
                    const double e = std::fabs(g / r - 1);
                    if (e > worst) worst = e;
                }
            }
            expect(worst <= 1e-10, "host build within 1e-10 of the oracle");
            std::printf("Z=%d mass=%g: %lld pairs, max rel %.2e\n", el.Z, mass, (long long)n, worst);
        }
    std::printf(failures ? "FAILED (%d)\n" : "ok\n", failures);
    return failures ? 1 : 0;
}
