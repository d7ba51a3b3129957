// Host-code driver for the sanitizer test: runs the XML reader and the output
// writers of libCReaDDyGPU (rd_xml.cpp, rd_output.cpp, built here with
// AddressSanitizer and UndefinedBehaviorSanitizer) on the reference's input
// fixtures, on malformed inputs and on edge-case frames.  The simulation
// entry points that rd_create_from_model calls are not part of this build;
// they are defined below to report failure, and are never reached.
#include "rdgpu.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

extern "C" {
int rd_create(const rd_params*, rd_system**) { return RD_EHIP; }
void rd_destroy(rd_system*) {}
const char* rd_last_error(const rd_system*) { return "host-only build"; }
int rd_set_particle_types(rd_system*, int, const double*, const double*, const double*) { return RD_EHIP; }
int rd_add_order_one_potential(rd_system*, const rd_pot1*) { return RD_EHIP; }
int rd_add_order_two_potential(rd_system*, const rd_pot2*, const uint8_t*) { return RD_EHIP; }
int rd_set_particles(rd_system*, int, const float*, const int*) { return RD_EHIP; }
int rd_set_bonds(rd_system*, int, const int*, const int*, int, const rd_pot2*, const double*) { return RD_EHIP; }
int rd_initialize(rd_system*) { return RD_EHIP; }
}

static int fails = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++fails;                                                    \
        }                                                               \
    } while (0)

static std::string P(const std::string& d, const char* f) { return d + "/" + f; }

static int read_dir(const std::string& d, const std::string& g, rd_model** m, char* err)
{
    std::string pp = P(d, "param_particles.xml"), tp = P(d, "tplgy_potentials.xml"), pg = P(g, "param_groups.xml");
    std::string tc = P(d, "tplgy_coordinates.xml"), tg = P(g, "tplgy_groups.xml"), gl = P(d, "param_global.xml");
    return rd_model_read_xml(pp.c_str(), tp.c_str(), pg.c_str(), tc.c_str(), tg.c_str(), gl.c_str(), m, err, 256);
}

int main(int argc, char** argv)
{
    if (argc < 4) return 2;
    const std::string fix = argv[1], groups = argv[2], tmp = argv[3];
    char err[256];
    // the reference's inputs with the test_default group files
    rd_model* m = nullptr;
    CHECK(read_dir(fix, groups, &m, err) == RD_OK);
    if (m) {
        rd_model_info I;
        CHECK(rd_model_get_info(m, &I) == RD_OK);
        CHECK(I.n_particles == 800 && I.n_types == 3 && I.n_bonds == 5 && I.n_bond_pots == 2);
        std::vector<double> D(I.n_types), rad(I.n_types), radii(I.n_types * I.n_types);
        CHECK(rd_model_types(m, D.data(), rad.data(), radii.data()) == RD_OK);
        std::vector<float> xyz(3 * I.n_particles);
        std::vector<int> type(I.n_particles), rid(I.n_particles);
        CHECK(rd_model_particles(m, xyz.data(), type.data(), rid.data()) == RD_OK);
        for (int i = 0; i < I.n_pot1; ++i) { rd_pot1 q; CHECK(rd_model_pot1(m, i, &q) == RD_OK); }
        std::vector<uint8_t> mask(I.n_types * I.n_types);
        for (int i = 0; i < I.n_pot2; ++i) { rd_pot2 q; CHECK(rd_model_pot2(m, i, &q, mask.data()) == RD_OK); }
        CHECK(rd_model_pot1(m, I.n_pot1, nullptr) != RD_OK);
        std::vector<int> bonds(2 * I.n_bonds), pot(I.n_bonds);
        std::vector<rd_pot2> bp(I.n_bond_pots);
        std::vector<double> r0(I.n_bond_pots);
        CHECK(rd_model_bonds(m, bonds.data(), pot.data(), bp.data(), r0.data()) == RD_OK);
        rd_system* h = nullptr;
        CHECK(rd_create_from_model(m, 10.0, 1, &h) != RD_OK && h == nullptr);
        rd_model_free(m);
    }
    // malformed inputs: a truncated global file, an unknown group template,
    // a missing coordinates file
    {
        std::string bad = tmp + "/bad";
        std::string cmd = "mkdir -p " + bad + " && cp " + fix + "/*.xml " + bad + "/";
        CHECK(system(cmd.c_str()) == 0);
        FILE* f = fopen((bad + "/param_global.xml").c_str(), "w");
        fputs("<param_global><T>293", f);
        fclose(f);
        rd_model* b = nullptr;
        CHECK(read_dir(bad, bad, &b, err) != RD_OK && b == nullptr);
        cmd = "cp " + fix + "/param_global.xml " + bad + "/";
        CHECK(system(cmd.c_str()) == 0);
        f = fopen((bad + "/tplgy_groups.xml").c_str(), "w");
        fputs("<tplgy_groups><g id=\"0\" type=\"7\" internalAndParticleId=\"[[0,0];[1,1]]\"/></tplgy_groups>", f);
        fclose(f);
        CHECK(read_dir(bad, bad, &b, err) != RD_OK && b == nullptr);
        f = fopen((bad + "/tplgy_groups.xml").c_str(), "w");
        fputs("<tplgy_groups><g id=\"0\" type=\"0\" internalAndParticleId=\"[[0,0];[1,", f);
        fclose(f);
        if (read_dir(bad, bad, &b, err) == RD_OK) rd_model_free(b);
        CHECK(rd_model_read_xml(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &b, err, 256) != RD_OK);
    }
    // writers: empty frame, special values, many particles across threads
    for (int fmt : { RD_TRAJ_PDB, RD_TRAJ_XML }) {
        rd_traj* t = nullptr;
        std::string path = tmp + (fmt == RD_TRAJ_PDB ? "/t.pdb" : "/t.xml");
        CHECK(rd_traj_open(path.c_str(), fmt, &t) == RD_OK);
        CHECK(rd_traj_write_frame(t, 0, 0, nullptr, nullptr) == RD_OK);
        const int n = 100000;
        std::vector<float> x(3 * n);
        std::vector<int> ty(n);
        for (int i = 0; i < 3 * n; ++i) x[i] = (float)((long long)i * 7919 % 200001 - 100000) * 0.01f;
        x[0] = NAN; x[1] = INFINITY; x[2] = -0.0f; x[3] = 1e30f;
        for (int i = 0; i < n; ++i) ty[i] = i % 7;
        CHECK(rd_traj_write_frame(t, 1, n, x.data(), ty.data()) == RD_OK);
        CHECK(rd_traj_close(t) == RD_OK);
    }
    rd_traj* t = nullptr;
    CHECK(rd_traj_open(nullptr, RD_TRAJ_PDB, &t) != RD_OK);
    std::vector<double> g(2 * 2 * 50, 1.0);
    const char* names[] = { "A", "B" };
    CHECK(rd_rdf_write((tmp + "/rdf.csv").c_str(), 2, names, 50, 10.0, g.data()) == RD_OK);
    CHECK(rd_rdf_write((tmp + "/rdf2.csv").c_str(), 2, nullptr, 50, 10.0, g.data()) == RD_OK);
    printf("host sanitizer run: %d failed checks\n", fails);
    return fails ? 1 : 0;
}
