// Spectral tables from files: the `.spd` reader and the two scene-builder entry points that turn
// the reference's tabulated optical constants (Source/Tables/ior/<name>.eta.spd, .k.spd) and
// emission spectra (Source/Tables/emission/<name>.spd) into spectral profiles
// (Utils/Color/Spectrum.cpp:89-112, Scene/SceneBuilder.cpp:911-1000).
//
// A `.spd` file is text: blank lines and lines starting with '#' (after trimming) are skipped;
// every other line holds "wavelength value" as two floats; the first line that does not parse ends
// the table. Profiles are built by plt_spectral_profile_build (PiecewiseLinearSpectrum ->
// SampledSpectrum -> genSpectralProfile).
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "plt_internal.hpp"

using plt::set_error;

namespace {

// Reads a `.spd` table; returns PLT_OK, or an error naming the file.
int read_spd(const std::string& path, std::vector<float>& wl, std::vector<float>& val) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return set_error(PLT_EINVAL, "Path could not be read '" + path + "'.");
    std::string line;
    wl.clear();
    val.clear();
    int c = 0;
    bool eof = false;
    while (!eof) {
        line.clear();
        while ((c = std::fgetc(f)) != EOF && c != '\n') line.push_back(char(c));
        eof = c == EOF;
        if (eof && line.empty()) break;
        size_t a = 0, b = line.size();
        while (a < b && std::isspace(static_cast<unsigned char>(line[a]))) ++a;
        while (b > a && std::isspace(static_cast<unsigned char>(line[b - 1]))) --b;
        if (a == b || line[a] == '#') continue;
        const std::string t = line.substr(a, b - a);
        const char* p = t.c_str();
        char* end = nullptr;
        errno = 0;
        const float lambda = std::strtof(p, &end);
        if (end == p) break;
        const char* q = end;
        const float value = std::strtof(q, &end);
        if (end == q) break;
        wl.push_back(lambda);
        val.push_back(value);
    }
    std::fclose(f);
    return PLT_OK;
}

int profile_from_table(const std::string& path, float scale, int compute_icdf, plt_spectral_profile* out) {
    std::vector<float> wl, val;
    if (int rc = read_spd(path, wl, val)) return rc;
    if (wl.size() < 2) return set_error(PLT_EINVAL, "spectral table '" + path + "' has fewer than two entries");
    if (scale != 1.f) {
        if (!(scale >= 0.f)) return set_error(PLT_EINVAL, "spectrum scale must be non-negative");
        for (float& v : val) v *= scale;  // PiecewiseLinearSpectrum::scale (Spectrum.cpp:114-119)
    }
    return plt_spectral_profile_build(uint32_t(wl.size()), wl.data(), val.data(), compute_icdf, out);
}

std::string join(const char* dir, const char* sub, const std::string& file) {
    std::string p = dir ? dir : "";
    if (!p.empty() && p.back() != '/') p.push_back('/');
    return p + sub + file;
}

}  // namespace

extern "C" {

int plt_spectrum_load_spd(const char* path, uint32_t capacity, float* wavelengths, float* values, uint32_t* count) {
    if (!path || !count) return set_error(PLT_EINVAL, "plt_spectrum_load_spd: bad arguments");
    std::vector<float> wl, val;
    if (int rc = read_spd(path, wl, val)) return rc;
    *count = uint32_t(wl.size());
    if (wl.size() > capacity) {
        if (capacity == 0 && !wavelengths && !values) return PLT_OK;  // size query
        return set_error(PLT_EINVAL, "plt_spectrum_load_spd: capacity below the table's " + std::to_string(wl.size()) +
                                         " entries");
    }
    if (wl.empty()) return PLT_OK;  // a table without entries: nothing to copy
    if (!wavelengths || !values) return set_error(PLT_EINVAL, "plt_spectrum_load_spd: NULL output");
    std::memcpy(wavelengths, wl.data(), wl.size() * sizeof(float));
    std::memcpy(values, val.data(), val.size() * sizeof(float));
    return PLT_OK;
}

// SampledSpectrum<float>(value): one sample over [380, 720] nm (SampledSpectrum.h:75-80), through
// genSpectralProfile without an inverse CDF: one bin holding the value.
int plt_spectral_profile_constant(float value, plt_spectral_profile* out) {
    if (!out) return set_error(PLT_EINVAL, "plt_spectral_profile_constant: out is NULL");
    std::memset(out, 0, sizeof(*out));
    out->intensity = 1.f;
    out->min_wavelength = 380.f;
    out->max_wavelength = 720.f;
    out->bins = 1;
    out->pdf[0] = value;
    return PLT_OK;
}

int plt_spectral_profiles_from_material(const char* tables_dir, const char* name, plt_spectral_profile* n_out,
                                        plt_spectral_profile* k_out) {
    if (!name || !n_out || !k_out) return set_error(PLT_EINVAL, "plt_spectral_profiles_from_material: bad arguments");
    if (std::strcmp(name, "none") == 0) {  // n = 1, k = 0 (SceneBuilder.cpp:974-982)
        plt_spectral_profile_constant(1.f, n_out);
        return plt_spectral_profile_constant(0.f, k_out);
    }
    if (int rc = profile_from_table(join(tables_dir, "ior/", std::string(name) + ".eta.spd"), 1.f, 0, n_out)) return rc;
    return profile_from_table(join(tables_dir, "ior/", std::string(name) + ".k.spd"), 1.f, 0, k_out);
}

int plt_spectral_profile_from_emitter(const char* tables_dir, const char* name, float scale, plt_spectral_profile* out) {
    if (!name || !out) return set_error(PLT_EINVAL, "plt_spectral_profile_from_emitter: bad arguments");
    const std::string path = join(tables_dir, "emission/", std::string(name) + ".spd");
    if (FILE* f = std::fopen(path.c_str(), "rb")) {
        std::fclose(f);
    } else {
        return set_error(PLT_EINVAL, "Emission profile '" + std::string(name) + "' not found.");
    }
    return profile_from_table(path, scale, 1, out);  // addSpectralProfile(..., computeICdf = true)
}

}  // extern "C"
