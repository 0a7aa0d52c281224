// Host-only driver for the AddressSanitizer / UndefinedBehaviorSanitizer build of the library's
// host code: the input-file readers on the reference's data files and on every file named on the
// command line (malformed inputs must fail with a status, never crash), and the FP64 table
// builder with PA_BATCH_HOST_ONLY (no device is touched) on valid and invalid batches.
//
// usage: driver <golden dir> [extra input files...]
#include <pureadsorb_amd.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

static int g_fail = 0;
#define EXPECT(cond)                                                             \
    do {                                                                         \
        if (!(cond)) {                                                           \
            char m_[1024];                                                       \
            pa_last_error(m_, sizeof m_);                                        \
            std::printf("FAILED %s:%d: %s (last error: %s)\n", __FILE__, __LINE__, #cond, m_); \
            ++g_fail;                                                            \
        }                                                                        \
    } while (0)

struct Inputs {
    std::vector<double> cell, frac, q, sigma, eps, sites, gq;
    std::vector<int32_t> types, gtypes;
    int32_t natoms = 0, ntypes = 0, tail = 0, nsites = 0;
    double cutoff = 0.0;
};

static bool read_all(const std::string &dir, Inputs &in)
{
    pa_ff_file *ff = nullptr;
    pa_cif_file *cif = nullptr;
    pa_guest_file *g = nullptr;
    if (pa_read_forcefield((dir + "/trappe.yaml").c_str(), &ff) != PA_OK) return false;
    if (pa_read_cif((dir + "/RUBTAK.cif").c_str(), &cif) != PA_OK) return false;
    if (pa_read_guest((dir + "/co2.yaml").c_str(), ff, &g) != PA_OK) return false;
    EXPECT(pa_ff_info(ff, &in.ntypes, &in.cutoff, &in.tail) == PA_OK);
    in.sigma.resize((size_t)in.ntypes * in.ntypes);
    in.eps.resize((size_t)in.ntypes * in.ntypes);
    std::vector<double> st(in.ntypes), et(in.ntypes);
    EXPECT(pa_ff_params(ff, st.data(), et.data(), in.sigma.data(), in.eps.data()) == PA_OK);
    size_t need = 0;
    EXPECT(pa_ff_names(ff, nullptr, 0, &need) == PA_OK);
    std::vector<char> names(need);
    EXPECT(pa_ff_names(ff, names.data(), names.size(), &need) == PA_OK);
    // a buffer one byte short must be refused, not overrun
    if (need > 1) EXPECT(pa_ff_names(ff, names.data(), need - 1, &need) != PA_OK);
    int32_t idx = -1;
    EXPECT(pa_ff_typeindex(ff, "C_co2", &idx) == PA_OK && idx >= 0);
    EXPECT(pa_ff_typeindex(ff, "no such type", &idx) != PA_OK);
    in.cell.resize(9);
    EXPECT(pa_cif_info(cif, &in.natoms, in.cell.data()) == PA_OK);
    in.frac.resize(3 * (size_t)in.natoms);
    in.q.resize(in.natoms);
    in.types.resize(in.natoms);
    EXPECT(pa_cif_atoms(cif, in.frac.data(), in.q.data()) == PA_OK);
    EXPECT(pa_cif_types(cif, ff, in.types.data()) == PA_OK);
    for (int which : {PA_CIF_LABELS, PA_CIF_SYMBOLS}) {
        EXPECT(pa_cif_strings(cif, which, nullptr, 0, &need) == PA_OK);
        std::vector<char> buf(need);
        EXPECT(pa_cif_strings(cif, which, buf.data(), buf.size(), &need) == PA_OK);
        if (need > 1) EXPECT(pa_cif_strings(cif, which, buf.data(), need - 1, &need) != PA_OK);
    }
    double tc, pc, om;
    EXPECT(pa_guest_info(g, &in.nsites, &tc, &pc, &om) == PA_OK);
    in.sites.resize(3 * (size_t)in.nsites);
    in.gtypes.resize(in.nsites);
    in.gq.resize(in.nsites);
    EXPECT(pa_guest_sites(g, in.sites.data(), in.gtypes.data(), in.gq.data()) == PA_OK);
    pa_cif_destroy(cif);
    pa_ff_destroy(ff);
    pa_guest_destroy(g);
    return true;
}

// replicate (src/structure.jl:79-92): image (i, j, k), i fastest, atoms inner
static void replicate(const Inputs &in, int r, std::vector<double> &cell, std::vector<double> &frac,
                      std::vector<double> &q, std::vector<int32_t> &types)
{
    cell = in.cell;
    for (int i = 0; i < 9; ++i) cell[i] *= r;
    frac.clear();
    q.clear();
    types.clear();
    for (int k = 0; k < r; ++k)
        for (int j = 0; j < r; ++j)
            for (int i = 0; i < r; ++i)
                for (int a = 0; a < in.natoms; ++a) {
                    frac.push_back((in.frac[3 * a] + i) / r);
                    frac.push_back((in.frac[3 * a + 1] + j) / r);
                    frac.push_back((in.frac[3 * a + 2] + k) / r);
                    q.push_back(in.q[a]);
                    types.push_back(in.types[a]);
                }
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::printf("usage: %s <golden dir> [files...]\n", argv[0]);
        return 2;
    }
    Inputs in;
    EXPECT(read_all(argv[1], in));
    EXPECT(in.natoms == 114 && in.nsites == 3);

    // malformed or foreign files: each reader returns a status and leaves no handle behind
    int refused = 0;
    for (int i = 2; i < argc; ++i) {
        pa_cif_file *cif = nullptr;
        pa_ff_file *ff = nullptr;
        if (pa_read_cif(argv[i], &cif) == PA_OK) pa_cif_destroy(cif); else { EXPECT(cif == nullptr); ++refused; }
        if (pa_read_forcefield(argv[i], &ff) == PA_OK) pa_ff_destroy(ff); else { EXPECT(ff == nullptr); ++refused; }
    }
    std::printf("malformed inputs: %d reader calls refused\n", refused);

    // the host-only table builder: RUBTAK 3x3x3 twice (deduplicated) + the unit cell once
    std::vector<double> c3, f3, q3, c2, f2, q2;
    std::vector<int32_t> t3, t2;
    replicate(in, 3, c3, f3, q3, t3);
    replicate(in, 1, c2, f2, q2, t2);
    pa_framework fws[3];
    for (int s = 0; s < 3; ++s) {
        const bool big = s != 1;
        std::memcpy(fws[s].cell, (big ? c3 : c2).data(), 9 * sizeof(double));
        fws[s].natoms = (int32_t)(big ? q3 : q2).size();
        fws[s].frac = (big ? f3 : f2).data();
        fws[s].types = (big ? t3 : t2).data();
        fws[s].charges = (big ? q3 : q2).data();
        for (int d = 0; d < 3; ++d) fws[s].replication[d] = big ? 3 : 1;
    }
    pa_forcefield ff = {in.ntypes, in.sigma.data(), in.eps.data(), in.cutoff, in.tail};
    pa_guest g = {in.nsites, in.sites.data(), in.gtypes.data(), in.gq.data()};
    pa_ewald ew = {12.0, 1e-6};
    pa_batch *b = nullptr;
    // the unit cell is narrower than 2 x 12 Å: an argument error naming the replication
    EXPECT(pa_batch_create(fws, 3, &ff, &g, &ew, PA_BATCH_HOST_ONLY, &b) == PA_ERR_ARG && b == nullptr);
    if (b) pa_batch_destroy(b);
    b = nullptr;
    for (uint32_t flags : {PA_BATCH_HOST_ONLY, PA_BATCH_HOST_ONLY | PA_BATCH_FULLK}) {
        pa_framework two[2] = {fws[0], fws[2]};
        EXPECT(pa_batch_create(two, 2, &ff, &g, &ew, flags, &b) == PA_OK && b);
        if (!b) continue;
        int32_t nfw = 0, nsys = 0;
        EXPECT(pa_batch_nframeworks(b, &nfw, &nsys) == PA_OK && nfw == 1 && nsys == 2);
        pa_framework_info fi;
        EXPECT(pa_batch_framework_info(b, 0, &fi) == PA_OK && fi.natoms == 3078);
        std::printf("host batch flags %u: %d atoms, %d of %d k-vectors kept, constant_offset %.12g eV\n", flags,
                    fi.natoms, fi.nk, fi.nk_full, fi.constant_offset);
        EXPECT(pa_batch_framework_info(b, 1, &fi) != PA_OK);
        pa_widom_result r[2];
        EXPECT(pa_widom(b, &g, 298.15, 1000, 1, 10, 0, 0, PA_PREC_MIXED, r) == PA_ERR_ARG);
        pa_batch_destroy(b);
        b = nullptr;
    }
    // invalid arguments
    std::vector<int32_t> bad = t3;
    bad[5] = in.ntypes + 3;
    pa_framework fb = fws[0];
    fb.types = bad.data();
    EXPECT(pa_batch_create(&fb, 1, &ff, &g, &ew, PA_BATCH_HOST_ONLY, &b) == PA_ERR_ARG && b == nullptr);
    fb = fws[0];
    fb.replication[1] = 0;
    EXPECT(pa_batch_create(&fb, 1, &ff, &g, &ew, PA_BATCH_HOST_ONLY, &b) == PA_ERR_ARG);
    EXPECT(pa_batch_create(fws, 0, &ff, &g, &ew, PA_BATCH_HOST_ONLY, &b) == PA_ERR_ARG);
    EXPECT(pa_batch_create(nullptr, 1, &ff, &g, &ew, PA_BATCH_HOST_ONLY, &b) == PA_ERR_ARG);
    pa_ewald tight = {12.0, 1e-30};
    EXPECT(pa_batch_create(fws, 1, &ff, &g, &tight, PA_BATCH_HOST_ONLY, &b) == PA_ERR_ARG);
    char msg[64];
    EXPECT(pa_last_error(msg, sizeof msg) > 0 && std::strlen(msg) < sizeof msg);

    std::printf(g_fail ? "SANITIZE_DRIVER_FAILED %d\n" : "SANITIZE_DRIVER_OK\n", g_fail);
    return g_fail ? 1 : 0;
}
