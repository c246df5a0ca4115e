/*
 * driver.cpp -- standalone driver of the density solve, linked against libcosmoSPHere.a (the C
 * slot transfers and the C++ solveDensH of include/cosmosphere.hpp, as the reference's driver
 * calls them, src/main.cu:117-141).  Command line and report follow the reference's
 * build/density_hip (src/main.cu:143-171):
 *
 *   density_hip lattice <npartx> [output_h_file | no_output] [jiggle]
 *   density_hip <datafile>       [output_h_file | no_output] [jiggle]
 *
 * A datafile is Fortran unformatted sequential (include/io.hpp:1-17): three records, each
 * framed by its 4-byte length before and after -- ngas (int32), pmass (float64), xyzh(4, ngas)
 * (float64, x y z h of particle 1, then of particle 2, ...).  "jiggle" scales h of about 1% of
 * the particles by 1 + 0.5 d, d = copysign(r^30, r - 0.5), r uniform (src/main.cu:48-63), so
 * that a converged dump needs several Newton iterations again.  The solve runs once to warm
 * up, then FLAT_PARTICLE and WARP_PER_LEAF from the same h; the second call's h is written one
 * value per line ("%.15e").
 *
 * Extensions: "mixed" selects the mixed precision form; "--check-read <datafile>" only reads
 * the file and prints its count, particle mass and sums (no device needed).
 */
#include "../../include/cosmosphere.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{

struct Particles
{
    int ngas = 0;
    double pmass = 0.0;
    std::vector<double> x, y, z, h;
};

// One record: start marker == expected payload size, payload, end marker == start marker.
void readRecord(std::FILE* fp, void* dst, size_t bytes, const char* what)
{
    int32_t head = 0, tail = 0;
    if (std::fread(&head, 4, 1, fp) != 1) throw std::runtime_error(std::string("truncated before the ") + what + " record");
    if ((size_t)(uint32_t)head != bytes)
        throw std::runtime_error(std::string("the ") + what + " record holds " + std::to_string(head) +
                                 " bytes, expected " + std::to_string(bytes));
    if (std::fread(dst, 1, bytes, fp) != bytes) throw std::runtime_error(std::string("truncated ") + what + " record");
    if (std::fread(&tail, 4, 1, fp) != 1 || tail != head)
        throw std::runtime_error(std::string("the ") + what + " record's end marker does not match its start");
}

Particles readDump(const std::string& path)
{
    std::FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) throw std::runtime_error("cannot open '" + path + "'");
    Particles p;
    try
    {
        int32_t n = 0;
        readRecord(fp, &n, 4, "ngas");
        if (n <= 0) throw std::runtime_error("ngas = " + std::to_string(n));
        readRecord(fp, &p.pmass, 8, "pmass");
        p.ngas = n;
        const size_t bytes = (size_t)4 * (size_t)n * 8;
        if (bytes > 0x7FFFFFFFu) throw std::runtime_error("xyzh record beyond a 32-bit length marker");
        int32_t head = 0, tail = 0;
        if (std::fread(&head, 4, 1, fp) != 1) throw std::runtime_error("truncated before the xyzh record");
        if ((size_t)(uint32_t)head != bytes)
            throw std::runtime_error("the xyzh record holds " + std::to_string(head) + " bytes, expected " +
                                     std::to_string(bytes));
        p.x.resize(n); p.y.resize(n); p.z.resize(n); p.h.resize(n);
        std::vector<double> chunk(4 * 32768);
        for (size_t done = 0; done < (size_t)n;)
        {
            const size_t m = std::min<size_t>(32768, (size_t)n - done);
            if (std::fread(chunk.data(), 8, 4 * m, fp) != 4 * m) throw std::runtime_error("truncated xyzh record");
            for (size_t k = 0; k < m; ++k)
            {
                p.x[done + k] = chunk[4 * k];
                p.y[done + k] = chunk[4 * k + 1];
                p.z[done + k] = chunk[4 * k + 2];
                p.h[done + k] = chunk[4 * k + 3];
            }
            done += m;
        }
        if (std::fread(&tail, 4, 1, fp) != 1 || tail != head)
            throw std::runtime_error("the xyzh record's end marker does not match its start");
    }
    catch (...)
    {
        std::fclose(fp);
        throw;
    }
    std::fclose(fp);
    return p;
}

// Close-packed lattice in the unit box (rows offset by half a spacing, layers by a third of a
// row), h = 1.2 x spacing, unit total mass: src/main.cu:84-109.
Particles lattice(int npartx)
{
    if (npartx < 4) throw std::runtime_error("npartx must be at least 4");
    const double dx = 1.0 / npartx, dy = dx * std::sqrt(3.0) / 2.0, dz = dx * std::sqrt(6.0) / 3.0;
    const int ny = (int)(1.0 / dy), nz = (int)(1.0 / dz);
    Particles p;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < npartx; ++i)
            {
                p.x.push_back(i * dx + ((j + k) % 2) * 0.5 * dx);
                p.y.push_back(j * dy + (k % 2) * dy / 3.0);
                p.z.push_back(k * dz);
            }
    p.ngas = (int)p.x.size();
    p.pmass = 1.0 / p.ngas;
    p.h.assign(p.ngas, 1.2 * dx);
    return p;
}

int jiggle(std::vector<double>& h)
{
    std::mt19937 gen(42);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    int count = 0;
    for (double& v : h)
    {
        if (u(gen) < 0.99) continue;
        const double r = u(gen);
        v *= 1.0 + 0.5 * std::copysign(std::pow(r, 30.0), r - 0.5);
        ++count;
    }
    return count;
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Range
{
    double lo, hi;
};
Range range(const std::vector<double>& v)
{
    const auto mm = std::minmax_element(v.begin(), v.end());
    return Range{*mm.first, *mm.second};
}

// The caller's four steps: size, send, solve, fetch.  Returns the iterations run.
int solve(const Particles& p, std::vector<double>& h, std::vector<double>& rho, int mode)
{
    const int n = p.ngas;
    cosmo_arrays_init(n);
    std::vector<double> pos(3 * (size_t)n);
    std::copy(p.x.begin(), p.x.end(), pos.begin());
    std::copy(p.y.begin(), p.y.end(), pos.begin() + n);
    std::copy(p.z.begin(), p.z.end(), pos.begin() + 2 * (size_t)n);
    cosmo_upload(COSMO_POS, pos.data(), n);
    cosmo_upload(COSMO_HSML, h.data(), n);
    const std::vector<double> zero(3 * (size_t)n, 0.0);
    cosmo_upload(COSMO_VEL, zero.data(), n);
    cosmo_upload(COSMO_ACCEL, zero.data(), n);
    const DensTimings t = solveDensH(n, p.pmass, mode == 1 ? KernelMode::WARP_PER_LEAF : KernelMode::FLAT_PARTICLE);
    const int iters = t.itersRun;
    cosmo_download(COSMO_HSML, h.data(), n);
    std::vector<double> dens(2 * (size_t)n);
    cosmo_download(COSMO_DENS_OUT, dens.data(), n);
    std::copy(dens.begin(), dens.begin() + n, rho.begin());
    return iters;
}

void report(const char* label, int iters, double wall, const std::vector<double>& h, const std::vector<double>& rho)
{
    std::printf("%s\n", label);
    std::printf(" Time bounding box (GPU)       :  %.4f s\n", cosmo_ext_get_stat("bbox"));
    std::printf(" Time Hilbert keys + GPU sort  :  %.4f s\n",
                cosmo_ext_get_stat("keys") + cosmo_ext_get_stat("sort") + cosmo_ext_get_stat("permute"));
    std::printf(" Time tree build (cs + linked) :  %.4f s\n", cosmo_ext_get_stat("tree"));
    std::printf(" Time node centres             :  %.4f s\n", cosmo_ext_get_stat("nodes"));
    std::printf(" Time j-leaf list build        :  %.4f s  (%d iter)\n", cosmo_ext_get_stat("jbuild"), iters);
    std::printf(" Time density kernel           :  %.4f s  (%d iter)\n", cosmo_ext_get_stat("nrkern"), iters);
    std::printf(" Total time building+solving   :  %.4f s\n", wall);
    const Range hr = range(h), rr = range(rho);
    std::printf(" Output h   range: [%.6e, %.6e]\n", hr.lo, hr.hi);
    std::printf(" Output rho range: [%.6e, %.6e]\n\n", rr.lo, rr.hi);
}

} // namespace

int main(int argc, char** argv)
{
    std::vector<std::string> pos;
    bool doJiggle = false, mixed = false;
    std::string checkRead;
    for (int i = 1; i < argc; ++i)
    {
        const std::string a = argv[i];
        if (a == "jiggle") doJiggle = true;
        else if (a == "mixed") mixed = true;
        else if (a == "--check-read" && i + 1 < argc) checkRead = argv[++i];
        else pos.push_back(a);
    }
    if (!checkRead.empty())
    {
        try
        {
            const Particles p = readDump(checkRead);
            double s[4] = {0, 0, 0, 0};
            for (int i = 0; i < p.ngas; ++i) { s[0] += p.x[i]; s[1] += p.y[i]; s[2] += p.z[i]; s[3] += p.h[i]; }
            std::printf("ngas %d pmass %.17g sums %.17g %.17g %.17g %.17g first %.17g %.17g %.17g %.17g\n", p.ngas,
                        p.pmass, s[0], s[1], s[2], s[3], p.x[0], p.y[0], p.z[0], p.h[0]);
            return 0;
        }
        catch (const std::exception& e)
        {
            std::fprintf(stderr, "Error reading file: %s\n", e.what());
            return 1;
        }
    }
    if (pos.empty())
    {
        std::fprintf(stderr, "Usage: %s lattice <npartx> [output_h_file] [jiggle] [mixed]\n", argv[0]);
        std::fprintf(stderr, "       %s <datafile> [output_h_file] [jiggle] [mixed]\n", argv[0]);
        std::fprintf(stderr, "       %s --check-read <datafile>\n", argv[0]);
        std::fprintf(stderr, "  output_h_file may be 'no_output' (default h_cpp.txt)\n");
        return 1;
    }
    const bool isLattice = pos[0] == "lattice";
    const int npartx = isLattice ? (pos.size() >= 2 ? std::atoi(pos[1].c_str()) : 50) : 0;
    const size_t outAt = isLattice ? 2 : 1;
    const std::string outFile = pos.size() > outAt ? pos[outAt] : "h_cpp.txt";

    std::printf("cosmoSPHere standalone density solver (gfx950)\n");
    const double t0 = now();
    Particles p;
    try
    {
        if (isLattice)
        {
            std::printf("Input     : close-packed lattice, npartx = %d\n", npartx);
            p = lattice(npartx);
        }
        else
        {
            std::printf("Data file : %s\n", pos[0].c_str());
            p = readDump(pos[0]);
        }
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "Error reading file: %s\n", e.what());
        return 1;
    }
    const Range h0 = range(p.h);
    std::printf("Finished read, ngas = %d\n pmass = %.6e\n maxh  = %.6e   minh = %.6e\n Time reading file: %.4f s\n",
                p.ngas, p.pmass, h0.hi, h0.lo, now() - t0);
    if (doJiggle)
    {
        const int nj = jiggle(p.h);
        const Range hj = range(p.h);
        std::printf("\n Jiggled %d particles (%.2f%%)\n maxh  = %.6e   minh = %.6e  (after jiggle)\n", nj,
                    100.0 * nj / p.ngas, hj.hi, hj.lo);
    }
    std::printf("\n");
    cosmo_ext_set_option("precision", mixed ? 1.0 : 0.0);

    std::vector<double> rho(p.ngas), h;
    std::printf("Warm-up call (results discarded)...\n");
    h = p.h;
    solve(p, h, rho, 0);
    std::printf("Warm-up complete.\n\n");

    double t = now();
    h = p.h;
    int it = solve(p, h, rho, 0);
    report("--- Call 1 (flat-particle) ---", it, now() - t, h, rho);

    t = now();
    h = p.h;
    it = solve(p, h, rho, 1);
    report("--- Call 2 (warp-per-leaf) ---", it, now() - t, h, rho);

    if (outFile != "no_output")
    {
        std::FILE* fp = std::fopen(outFile.c_str(), "w");
        if (!fp)
        {
            std::fprintf(stderr, "Warning: could not write output: cannot open '%s'\n", outFile.c_str());
            return 0;
        }
        for (double v : h) std::fprintf(fp, "%.15e\n", v);
        std::fclose(fp);
        std::printf(" h written to: %s\n", outFile.c_str());
    }
    return 0;
}
