/*
 * Randoms.h -- host-side random numbers for example set-up code, with the
 * reference's class name and interface (include/Randoms.h): Randoms(seed),
 * Uniforme() in (0, 1), Normal(avg, sigma), static sorte(m) in [-m, m].
 *
 * Uniforme() draws from the Park-Miller "minimal standard" generator
 * (multiplier 16807, modulus 2^31-1; std::minstd_rand0), Normal() by the
 * Box-Muller transform; the sequences are not the reference's.
 */
#pragma once

#include <cmath>
#include <cstdlib>
#include <random>

class Randoms {
public:
    explicit Randoms(long x) : gen_((unsigned long)(x > 0 ? x : (x < 0 ? -x : 1))) {}

    // uniform in (0, 1)
    double Uniforme() {
        const double m = (double)std::minstd_rand0::modulus;
        return (double)gen_() / m;  // gen_() lies in [1, m-1]
    }
    double Normal(double avg, double sigma) { return avg + sigma * gauss(); }
    // uniform in [-m, m] from the C library generator
    static double sorte(int m) { return (1.0 * std::rand()) / (1.0 * RAND_MAX) * 2.0 * m - m; }

private:
    std::minstd_rand0 gen_;
    bool have_spare_ = false;
    double spare_ = 0.0;

    double gauss() {
        if (have_spare_) {
            have_spare_ = false;
            return spare_;
        }
        const double u1 = Uniforme(), u2 = Uniforme();
        const double rad = std::sqrt(-2.0 * std::log(u1));
        const double ang = 6.283185307179586 * u2;
        spare_ = rad * std::sin(ang);
        have_spare_ = true;
        return rad * std::cos(ang);
    }
};
