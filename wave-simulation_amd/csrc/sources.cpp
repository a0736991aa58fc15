// Synthetic source signals of the reference's source files (wavelet type 1):
// Sources::generateSyntheticSignal (src/Acquisition/Sources.cpp:157-220) with the
// shapes of src/Acquisition/SourceSignal/*.cpp, evaluated in the reference's value
// type (float) step by step as there, so the signals equal the reference float
// build's.  Host code only; no GPU and no handle needed.
#include <cmath>
#include <string>
#include <vector>

#include "handle.hpp"

namespace {

float fsign(float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f); }

// [i1, i2] of the shapes that are non-zero for one period after Tshift, clamped to the signal
bool window(int NT, float DT, float FC, float ts, long &i1, long &i2)
{
    i1 = (long)std::floor(ts / DT);
    i2 = i1 + (long)std::floor(1.0 / FC / DT);
    if (i1 < 0 || i1 >= NT) return false;
    if (i2 >= NT) i2 = NT - 1;
    return true;
}

}  // namespace

extern "C" int wv_source_signal(int shape, int NT, double DT_, double FC_, double AMP_, double TS_, float *out)
{
    if (NT <= 0) return fail(WV_EINVAL, "NT is < 0: No valid argument!");
    if (!out) return fail(WV_EINVAL, "null output");
    const float DT = (float)DT_, FC = (float)FC_, AMP = (float)AMP_, TS = (float)TS_;
    if (!(DT > 0)) return fail(WV_EINVAL, "DT is < 0: No valid argument!");
    if (shape != 5 && !(FC > 0)) return fail(WV_EINVAL, "FC is < 0: No valid argument!");
    std::vector<float> s(NT, 0.f);
    auto tk = [&](int k) { return 0.f + (float)k * DT; };  // linearDenseVector(NT, 0, DT)
    switch (shape) {
    case 1: {  // Ricker (Ricker.cpp:29-53)
        const float help = (float)(1.5 / FC + TS);
        const float pf = (float)(M_PI * FC);
        for (int k = 0; k < NT; k++) {
            float tau = tk(k) - help;
            tau *= pf;
            const float h = tau * tau;
            const float e = std::exp(-1.f * h);
            const float w = 1.f - 2.f * h;
            s[k] = AMP * w * e;
        }
        break;
    }
    case 2: {  // SinW (SinW.cpp): sin(tau) - sin(2 tau) / 2 over one period after Tshift
        long i1, i2;
        if (!window(NT, DT, FC, TS, i1, i2)) return fail(WV_EINVAL, "Signal is shifted to Tshift > Tsimulation");
        for (long i = i1, c = 0; i <= i2; i++, c++) {
            const double th = 2.0 * c * DT * M_PI * FC;
            const float a = (float)std::sin(th), b = (float)std::sin(2.0 * th);
            s[i] = AMP * (a - b / 2.f);
        }
        break;
    }
    case 3: {  // SinThree (SinThree.cpp): sin^3(pi FC (t - Tshift)) over one period
        long i1, i2;
        if (!window(NT, DT, FC, TS, i1, i2)) return fail(WV_EINVAL, "Signal is shifted to Tshift > Tsimulation");
        for (long i = i1, c = 0; i <= i2; i++, c++) {
            const double v = std::pow(std::sin(c * DT * M_PI * FC), 3);
            s[i] = AMP * (float)v;
        }
        break;
    }
    case 4: {  // FGaussian (FGaussian.cpp): first derivative of a Gaussian, centre 1.2/FC + Tshift
        const float help = (float)(1.2 / FC + TS);
        const float pf = (float)(M_PI * FC);
        for (int k = 0; k < NT; k++) {
            float tau = tk(k) - help;
            tau *= pf;
            const float h = -2.f * tau;
            const float e = std::exp(-1.f * tau * tau);
            s[k] = AMP * h * e;
        }
        break;
    }
    case 5: {  // Spike (Spike.cpp): AMP at floor(Tshift / DT)
        const long i = (long)std::floor(TS / DT);
        if (i < 0 || i >= NT) return fail(WV_EINVAL, "Signal is shifted to Tshift > Tsimulation");
        s[i] = AMP * 1.f;
        break;
    }
    case 6: {  // IntgSinThree (IntgSinThree.cpp): (0.5 - 0.75 cos + 0.25 cos^3) / FC / pi / 0.75; the
               // constant 0.5 of the formula is added over the whole signal, as there
        long i1, i2;
        if (!window(NT, DT, FC, TS, i1, i2)) return fail(WV_EINVAL, "Signal is shifted to Tshift > Tsimulation");
        std::vector<float> help(NT, 0.f), zero(NT, 0.f);
        for (long i = i1, c = 0; i <= i2; i++, c++) {
            const double cs = std::cos(c * DT * M_PI * FC);
            help[i] = (float)cs;
            zero[i] = (float)std::pow(cs, 3);
        }
        for (int k = 0; k < NT; k++) {
            float z = 0.25f * zero[k];
            z = z - 0.75f * help[k];
            z = 0.5f + z;
            z = z / FC;
            z = z / (float)M_PI;
            z = z / 0.75f;
            s[k] = AMP * z;
        }
        break;
    }
    case 7: {  // Ricker_GprMax (Ricker_GprMax.cpp): -2 zeta exp(1/(4 zeta)) exp(-zeta tau^2) tau, tau = t - 1/FC - Tshift
        const float help0 = 1.f / FC + TS;
        const float zeta = (float)(2.0 * M_PI * M_PI * FC * FC);
        float z = zeta * std::exp((1.f / zeta) / 4.f);
        z = -2.f * z;
        for (int k = 0; k < NT; k++) {
            const float tau = tk(k) - help0;
            float h = tau * tau;
            h = zeta * h;
            h = -1.f * h;
            h = std::exp(h);
            s[k] = AMP * z * (h * tau);
        }
        break;
    }
    case 8: {  // Berlage (Berlage.cpp; Aldridge 1990): cos(2 pi FC tau) tau^2 exp(-2 FC tau) for tau >= 0,
               // tau = t - 1/FC - Tshift, scaled to AMP at its largest magnitude
        const float help0 = (float)(1.0 / FC + TS);
        const float alpha = 2 * FC;
        float mx = 0.f;
        for (int k = 0; k < NT; k++) {
            const float tau = tk(k) - help0;
            const float c = std::cos((float)(2 * M_PI * FC) * tau);
            const float hv = fsign(fsign(tau) + 1.f);
            float v = c * std::pow(tau, 2.f);
            v *= std::exp(tau * -alpha);
            v *= hv;
            s[k] = v;
            mx = std::fmax(mx, std::fabs(v));
        }
        if (mx > 0.f)
            for (auto &v : s) v *= AMP / mx;
        break;
    }
    case 9: {  // Sin (Sin.cpp): AMP sin(2 pi FC (t - Tshift))
        const float w = (float)(2.0 * M_PI * FC);
        for (int k = 0; k < NT; k++) {
            float tau = tk(k);
            tau -= TS;
            tau *= w;
            s[k] = AMP * std::sin(tau);
        }
        break;
    }
    default:
        return fail(WV_EINVAL, "Unknown wavelet shape ");
    }
    std::copy(s.begin(), s.end(), out);
    return WV_OK;
}
