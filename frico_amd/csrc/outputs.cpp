// Text outputs of the analyses, byte for byte as the reference writes them with std::println:
//   port_sweep.txt       write_sweep_data            (src/emw_ana_antenna.cpp:150-169)
//   polar_<n>.txt        write_radiation_diagrams    (src/emw_ana_antenna.cpp:171-203)
//   bistatic_rcs_<n>.txt postpro_context (bistatic)  (src/emw_ana_bistatic_rcs.cpp:97-117)
// std::format's "{}" for a double is std::to_chars(first, last, value): the shortest string that
// reads back to the same value, scientific or fixed, whichever is shorter; "{:.1f}" is
// std::to_chars(first, last, value, std::chars_format::fixed, 1).  The writers below use exactly
// those calls.
#include "../../include/frico_gpu.h"
#include "common.hpp"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <string>

namespace frico {

static void put(std::string& s, double v)
{
    char b[64];
    const auto r = std::to_chars(b, b + sizeof b, v);
    s.append(b, r.ptr);
}

static void put_fixed(std::string& s, double v, int prec)
{
    char b[512];
    const auto r = std::to_chars(b, b + sizeof b, v, std::chars_format::fixed, prec);
    s.append(b, r.ptr);
}

static void put(std::string& s, long long v)
{
    char b[32];
    const auto r = std::to_chars(b, b + sizeof b, v);
    s.append(b, r.ptr);
}

static int write_file(const char* path, const std::string& text)
{
    std::FILE* f = std::fopen(path, "wb");
    if (!f)
        return set_error("Can't open '%s' for writing", path), FRICO_ERR_STATE;
    const size_t n = std::fwrite(text.data(), 1, text.size(), f);
    const bool ok = n == text.size() && std::fclose(f) == 0;
    if (!ok)
        return set_error("writing '%s' failed", path), FRICO_ERR_STATE;
    return 0;
}

} // namespace frico

extern "C" {

int frico_format_double(double v, char* buf, int64_t size)
{
    using namespace frico;
    if (!buf || size < 1)
        return -2;
    std::string s;
    put(s, v);
    if ((int64_t)s.size() >= size)
        return set_error("buffer of %lld bytes too small", (long long)size), -3;
    std::snprintf(buf, (size_t)size, "%s", s.c_str());
    return 0;
}

int frico_write_port_sweep(const char* path, double Z0, int64_t nsteps, const double* freqs, const double* zin,
                           const double* swr)
{
    using namespace frico;
    if (!path)
        return set_error("null path"), -1;
    if (nsteps < 0 || (nsteps > 0 && (!freqs || !zin || !swr)))
        return set_error("null sweep arrays"), -3;
    std::string s = "# Port sweep results, Z0 = ";
    put_fixed(s, Z0, 1);
    s += " Ohm\n# step  frequency    Re(Z)    Im(Z)    SWR\n";
    for (int64_t i = 0; i < nsteps; i++) {
        put(s, (long long)i);
        s += ' ';
        put(s, freqs[i]);
        s += ' ';
        put(s, zin[2 * i]);
        s += ' ';
        put(s, zin[2 * i + 1]);
        s += ' ';
        put(s, swr[i]);
        s += '\n';
    }
    return write_file(path, s);
}

int frico_write_radiation_diagram(const char* path, double freq, const double* gain)
{
    using namespace frico;
    if (!path)
        return set_error("null path"), -1;
    if (!gain)
        return set_error("null gain"), -3;
    std::string s = "# FRICO radiation patterns. Frequency = ";
    put(s, freq);
    s += "\n# angle XY_lin XY_dB YZ_lin YZ_dB XZ_lin XZ_dB\n";
    for (int i = 0; i < 360; i++) {
        put(s, (long long)i);
        for (int p = 0; p < 3; p++) {
            const double g = gain[360 * p + i];
            s += ' ';
            put(s, g);
            s += ' ';
            put(s, 10.0 * std::log10(g));
        }
        s += '\n';
    }
    return write_file(path, s);
}

int frico_write_bistatic_rcs(const char* path, const double* table)
{
    using namespace frico;
    if (!path)
        return set_error("null path"), -1;
    if (!table)
        return set_error("null table"), -2;
    std::string s = "# FRICO bistatic RCS computation\n# Angle, VV_theta, HH_theta, VV_phi, HH_phi\n";
    for (int i = 0; i < 359; i++) {
        put(s, (long long)i);
        for (int c = 0; c < 4; c++) {
            s += ' ';
            put(s, table[4 * i + c]);
        }
        s += '\n';
    }
    return write_file(path, s);
}

} // extern "C"
