/*
 * Host-side producers of the formal solver's inputs.  They run once per
 * model (or per statistical-equilibrium iteration) on the CPU, outside the
 * device solve, and follow the reference's producers:
 *   Lines::set_inverse_width             src/model/lines/lines.tpp:45-60
 *   LineProducingSpecies::update_using_LTE
 *                                        src/model/lines/lineProducingSpecies/lineProducingSpecies.tpp:52-75
 *   get_emissivity / get_opacity         lineProducingSpecies.tpp:23-44
 *   Lines::set_emissivity_and_opacity    src/model/lines/lines.tpp:29-41
 *   Model::compute_spectral_discretisation  src/model/model.cpp:70-161, 257-298
 *   Model::set_dshift_max                src/model/model.cpp:1158-1178
 *   InterpHelper::InterpHelper           src/solver/solver_interp_helper.hpp:15-30
 *   Lines::read (sorted lines), heapsort src/model/lines/lines.cpp:44-64,
 *                                        src/tools/heapsort.hpp:64-88
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "mgx.h"
#include "mgx_constants.h"
#include "mgx_error.h"

namespace {

/* Thermodynamics::profile_width(inverse_mass, p) (thermodynamics.tpp:32-35) */
inline double profile_width(double inverse_mass, double T, double vturb2)
{
    return std::sqrt(MGX_TWO_KB_OVER_AMU_CC_SQUARED * inverse_mass * T + vturb2);
}

uint32_t total_lines(uint32_t nlspecs, const mgx_linedata *ld)
{
    uint32_t L = 0;
    for (uint32_t s = 0; s < nlspecs; s++) L += ld[s].nrad;
    return L;
}

int check_ld(uint32_t nlspecs, const mgx_linedata *ld)
{
    if (nlspecs == 0 || !ld) return mgx::fail(MGX_E_ARG, "no line-producing species");
    for (uint32_t s = 0; s < nlspecs; s++)
        if (ld[s].nrad && (!ld[s].frequency || !ld[s].irad || !ld[s].jrad))
            return mgx::fail(MGX_E_ARG, "incomplete line data for species %u", s);
    return MGX_OK;
}

/* tools/heapsort.hpp:8-88 (same sift order, so equal frequencies keep the reference's tie order) */
long hs_max(const std::vector<double> &a, long n, long i, long j, long k)
{
    long m = i;
    if (j < n && a[j] > a[m]) m = j;
    if (k < n && a[k] > a[m]) m = k;
    return m;
}

void hs_down(std::vector<double> &a, std::vector<uint32_t> &b, long n, long i)
{
    for (;;) {
        const long j = hs_max(a, n, i, 2 * i + 1, 2 * i + 2);
        if (j == i) break;
        std::swap(a[i], a[j]);
        std::swap(b[i], b[j]);
        i = j;
    }
}

} // namespace

extern "C" {

int mgx_compute_inverse_widths(uint32_t P, uint32_t nlspecs, const mgx_linedata *ld,
                               const double *T, const double *vturb2, double *inverse_width)
{
    mgx::clear_error();
    if (int e = check_ld(nlspecs, ld)) return e;
    if (!T || !vturb2 || !inverse_width) return mgx::fail(MGX_E_ARG, "null array");
    const uint32_t L = total_lines(nlspecs, ld);
#pragma omp parallel for schedule(static)
    for (uint32_t p = 0; p < P; p++) {
        uint32_t lid = 0;
        for (uint32_t s = 0; s < nlspecs; s++)
            for (uint32_t k = 0; k < ld[s].nrad; k++, lid++) {
                const double width = ld[s].frequency[k] * profile_width(ld[s].inverse_mass, T[p], vturb2[p]);
                inverse_width[(size_t)p * L + lid] = 1.0 / width;
            }
    }
    return MGX_OK;
}

int mgx_compute_lte(uint32_t P, uint32_t nspec, const double *abundance, const double *T,
                    uint32_t nlspecs, const mgx_linedata *ld, double min_line_emissivity,
                    double min_line_opacity, double *population, double *population_tot,
                    double *emissivity, double *opacity)
{
    mgx::clear_error();
    if (int e = check_ld(nlspecs, ld)) return e;
    if (!abundance || !T || !population || !population_tot || !emissivity || !opacity)
        return mgx::fail(MGX_E_ARG, "null array");
    for (uint32_t s = 0; s < nlspecs; s++)
        if (ld[s].num >= nspec) return mgx::fail(MGX_E_ARG, "species number out of range");
    const uint32_t L = total_lines(nlspecs, ld);
    size_t off = 0;
    uint32_t lid0 = 0;
    for (uint32_t s = 0; s < nlspecs; s++) {
        const mgx_linedata &d = ld[s];
        double *pop_s = population + off;
#pragma omp parallel for schedule(static)
        for (uint32_t p = 0; p < P; p++) {
            std::vector<long double> pop(d.nlev);
            const double ptot = abundance[(size_t)p * nspec + d.num];
            population_tot[(size_t)s * P + p] = ptot;
            double partition_function = 0.0;
            for (uint32_t i = 0; i < d.nlev; i++) {
                pop[i] = d.weight[i] * std::exp(-d.energy[i] / (MGX_KB * T[p]));
                partition_function += pop[i];
            }
            for (uint32_t i = 0; i < d.nlev; i++) pop[i] *= ptot / partition_function;
            for (uint32_t i = 0; i < d.nlev; i++) pop_s[(size_t)p * d.nlev + i] = (double)pop[i];
            for (uint32_t k = 0; k < d.nrad; k++) {
                const long double ni = pop[d.irad[k]];
                const long double nj = pop[d.jrad[k]];
                const double freq = d.frequency[k];
                const long double eps = freq * MGX_HH_OVER_FOUR_PI * d.A[k] * ni;
                const long double kap = freq * MGX_HH_OVER_FOUR_PI * (nj * d.Ba[k] - ni * d.Bs[k]);
                const size_t idx = (size_t)p * L + lid0 + k;
                emissivity[idx] = (double)std::max(eps, (long double)min_line_emissivity);
                opacity[idx] = (double)std::max(kap, (long double)min_line_opacity);
            }
        }
        off += (size_t)P * d.nlev;
        lid0 += d.nrad;
    }
    return MGX_OK;
}

int mgx_compute_spectral_discretisation(uint32_t P, uint32_t nlspecs, const mgx_linedata *ld,
                                        uint32_t nquads, const double *roots, const double *T,
                                        const double *vturb2, double *nu, uint32_t *f_line,
                                        uint32_t *f_species, uint32_t *f_k, uint32_t *f_z,
                                        uint8_t *f_in_line)
{
    mgx::clear_error();
    if (int e = check_ld(nlspecs, ld)) return e;
    if (!roots || !T || !vturb2 || !nu || !f_line || !f_species || !f_k || !f_z || !f_in_line)
        return mgx::fail(MGX_E_ARG, "null array");
    const uint32_t L = total_lines(nlspecs, ld);
    const uint32_t F = L * nquads;
#pragma omp parallel for schedule(static)
    for (uint32_t p = 0; p < P; p++) {
        uint32_t idx = 0;
        for (uint32_t s = 0; s < nlspecs; s++) {
            const double inverse_mass = ld[s].inverse_mass;
            for (uint32_t k = 0; k < ld[s].nrad; k++) {
                const double freqs_line = ld[s].frequency[k];
                const double width = freqs_line * profile_width(inverse_mass, T[p], vturb2[p]);
                for (uint32_t z = 0; z < nquads; z++)
                    nu[(size_t)p * F + idx++] = freqs_line + width * roots[(size_t)s * nquads + z];
            }
        }
    }
    uint32_t idx = 0, lid = 0;
    for (uint32_t s = 0; s < nlspecs; s++)
        for (uint32_t k = 0; k < ld[s].nrad; k++, lid++)
            for (uint32_t z = 0; z < nquads; z++, idx++) {
                f_line[idx] = lid;
                f_species[idx] = s;
                f_k[idx] = k;
                f_z[idx] = z;
                f_in_line[idx] = 1;
            }
    return MGX_OK;
}

int mgx_compute_image_discretisation(uint32_t P, double nu_min, double nu_max, uint32_t n,
                                     double *nu, uint32_t *f_line, uint32_t *f_species,
                                     uint32_t *f_k, uint32_t *f_z, uint8_t *f_in_line)
{
    mgx::clear_error();
    if (n < 1)
        return mgx::fail(MGX_E_ARG, "At least a single frequency is needed to compute a "
                                    "spectral discretization.");
    if (!nu || !f_line || !f_species || !f_k || !f_z || !f_in_line)
        return mgx::fail(MGX_E_ARG, "null array");
    if (n == 1) {
        for (uint32_t p = 0; p < P; p++) nu[p] = (double)(nu_min + nu_max) / 2.0;
    } else {
        const long double dnu = (nu_max - nu_min) / (n - 1);
#pragma omp parallel for schedule(static)
        for (uint32_t p = 0; p < P; p++)
            for (uint32_t f = 0; f < n; f++) nu[(size_t)p * n + f] = (double)(nu_min + f * dnu);
    }
    for (uint32_t f = 0; f < n; f++) {
        f_in_line[f] = 0;
        f_line[f] = 0; /* unused by the None and CloseLines imagers */
        f_species[f] = n;
        f_k[f] = n;
        f_z[f] = n;
    }
    return MGX_OK;
}

int mgx_compute_dshift_max(uint32_t P, uint32_t nlspecs, const mgx_linedata *ld, const double *T,
                           const double *vturb2, double max_width_fraction, double *dshift_max)
{
    mgx::clear_error();
    if (int e = check_ld(nlspecs, ld)) return e;
    if (!T || !vturb2 || !dshift_max) return mgx::fail(MGX_E_ARG, "null array");
#pragma omp parallel for schedule(static)
    for (uint32_t o = 0; o < P; o++) {
        double m = std::numeric_limits<double>::max();
        for (uint32_t s = 0; s < nlspecs; s++) {
            const double v = max_width_fraction * profile_width(ld[s].inverse_mass, T[o], vturb2[o]);
            if (m > v) m = v;
        }
        dshift_max[o] = m;
    }
    return MGX_OK;
}

int mgx_compute_interp_crit(uint32_t P, uint32_t nlspecs, const mgx_linedata *ld,
                            const double *opacity, const double *population_tot,
                            double *interp_crit)
{
    mgx::clear_error();
    if (int e = check_ld(nlspecs, ld)) return e;
    if (!opacity || !population_tot || !interp_crit) return mgx::fail(MGX_E_ARG, "null array");
    const uint32_t L = total_lines(nlspecs, ld);
#pragma omp parallel for schedule(static)
    for (uint32_t p = 0; p < P; p++) {
        uint32_t lid = 0;
        for (uint32_t s = 0; s < nlspecs; s++)
            for (uint32_t k = 0; k < ld[s].nrad; k++, lid++)
                interp_crit[(size_t)p * L + lid] =
                    opacity[(size_t)p * L + lid] / population_tot[(size_t)s * P + p];
    }
    return MGX_OK;
}

int mgx_sort_lines(uint32_t nlines, const double *line, double *sorted_line,
                   uint32_t *sorted_line_map)
{
    mgx::clear_error();
    if (!line || !sorted_line || !sorted_line_map) return mgx::fail(MGX_E_ARG, "null array");
    std::vector<double> a(line, line + nlines);
    std::vector<uint32_t> b(nlines);
    for (uint32_t i = 0; i < nlines; i++) b[i] = i;
    const long n = (long)nlines;
    for (long i = (n - 2) / 2; i >= 0; i--) hs_down(a, b, n, i);
    for (long i = 0; i < n; i++) {
        std::swap(a[n - i - 1], a[0]);
        std::swap(b[n - i - 1], b[0]);
        hs_down(a, b, n - i - 1, 0);
    }
    std::memcpy(sorted_line, a.data(), nlines * sizeof(double));
    std::memcpy(sorted_line_map, b.data(), nlines * sizeof(uint32_t));
    return MGX_OK;
}

} // extern "C"
