// rd_output.cpp -- trajectory and RDF writers of the reference's stand-alone
// driver (lib/ReaDDyGPU_Cpp/outputHandler/outputHandler.cpp):
//
//   PDB   PDBTrajectoryOutputHandler::writeFrame (outputHandler.cpp:79-103):
//         "MODEL     <frame>", one ATOM card per particle, "ENDMDL"
//   XML   XMLTrajectoryOutputHandler (outputHandler.cpp:110-148): <traj>, one
//         <tplgy_coords stepId> block per frame with <p id type c/> lines
//   RDF   RDFOutputHandler (outputHandler.cpp:17-58): header
//         "d,\tRDF(a,b)..." then one row per bin
//
// The text is the reference's printf formats, byte for byte, except the XML
// closing tag, which the reference writes as "<" TAB "raj>" (an escape slip in
// "<\traj>"); this writer closes with "</traj>".  Frames of millions of
// particles are formatted in parallel by a fixed-point formatter that
// reproduces printf's "%W.3f" exactly (a float times 1000 is exact in double,
// rounded half to even as printf does), then written in order.
#include "rdgpu.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

// printf("%*.3f", width, (double)v)
void put_f3(std::string& o, float v, int width)
{
    const double d = (double)v;
    const double s = d * 1000.0;                     // exact: 24 + 10 bits
    if (!std::isfinite(d) || std::fabs(s) >= 9.0e15) {
        char b[400];
        snprintf(b, sizeof(b), "%*.3f", width, d);
        o += b;
        return;
    }
    const double r = std::nearbyint(s);              // round half to even
    unsigned long long a = (unsigned long long)std::fabs(r);
    char dig[32];
    int nd = 0;
    unsigned long long ip = a / 1000ull;
    const unsigned frac = (unsigned)(a % 1000ull);
    do {
        dig[nd++] = (char)('0' + ip % 10ull);
        ip /= 10ull;
    } while (ip);
    const bool neg = std::signbit(d);
    const int len = (neg ? 1 : 0) + nd + 4;
    for (int k = len; k < width; ++k) o += ' ';
    if (neg) o += '-';
    while (nd) o += dig[--nd];
    o += '.';
    o += (char)('0' + frac / 100);
    o += (char)('0' + frac / 10 % 10);
    o += (char)('0' + frac % 10);
}

// printf("%*d", width, v)
void put_i(std::string& o, long long v, int width)
{
    char dig[24];
    int nd = 0;
    unsigned long long a = v < 0 ? 0ull - (unsigned long long)v : (unsigned long long)v;
    do {
        dig[nd++] = (char)('0' + a % 10ull);
        a /= 10ull;
    } while (a);
    const int len = nd + (v < 0 ? 1 : 0);
    for (int k = len; k < width; ++k) o += ' ';
    if (v < 0) o += '-';
    while (nd) o += dig[--nd];
}

void pdb_lines(std::string& o, int a0, int a1, const float* xyz, const int* type)
{
    for (int a = a0; a < a1; ++a) {
        o += "ATOM  ";
        put_i(o, a + 1, 5);
        o += "   ";
        put_i(o, type[a], 0);
        o += "    ";
        put_i(o, type[a], 0);
        o += "     1    ";
        put_f3(o, xyz[3 * a], 8);
        put_f3(o, xyz[3 * a + 1], 8);
        put_f3(o, xyz[3 * a + 2], 8);
        o += " 15.00  0.00\n";
    }
}

void xml_lines(std::string& o, int a0, int a1, const float* xyz, const int* type)
{
    for (int a = a0; a < a1; ++a) {
        o += "<p id=\"";
        put_i(o, a + 1, 0);
        o += "\" type=\"";
        put_i(o, type[a], 0);
        o += "\" c=\"";
        put_f3(o, xyz[3 * a], 58);
        o += ',';
        put_f3(o, xyz[3 * a + 1], 58);
        o += ',';
        put_f3(o, xyz[3 * a + 2], 58);
        o += "\"/>\n";
    }
}

}  // namespace

struct rd_traj {
    FILE* f = nullptr;
    int format = 0;
};

extern "C" int rd_traj_open(const char* path, int format, rd_traj** out)
{
    if (!path || !out || (format != RD_TRAJ_PDB && format != RD_TRAJ_XML)) return RD_EINVAL;
    *out = nullptr;
    FILE* f = fopen(path, "w");
    if (!f) return RD_EINVAL;
    if (format == RD_TRAJ_XML && fputs("<traj>\n", f) < 0) {
        fclose(f);
        return RD_EINVAL;
    }
    rd_traj* t = new rd_traj();
    t->f = f;
    t->format = format;
    *out = t;
    return RD_OK;
}

extern "C" int rd_traj_write_frame(rd_traj* t, int frame, int n, const float* xyz, const int* type)
{
    if (!t || n < 0 || (n > 0 && (!xyz || !type))) return RD_EINVAL;
    std::string head, tail;
    if (t->format == RD_TRAJ_PDB) {
        head = "MODEL     " + std::to_string(frame) + "\n";
        tail = "ENDMDL\n";
    } else {
        head = "<tplgy_coords version=\"1.1\" stepId=\"" + std::to_string(frame) + "\">\n";
        tail = "</tplgy_coords>\n";
    }
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nchunk = (int)std::max(1L, std::min<long>(std::min<unsigned>(hw, 16u), n / 20000));
    std::vector<std::string> buf((size_t)nchunk);
    auto work = [&](int c) {
        const int a0 = (int)((long)n * c / nchunk), a1 = (int)((long)n * (c + 1) / nchunk);
        buf[c].reserve((size_t)(a1 - a0) * (t->format == RD_TRAJ_PDB ? 66 : 200));
        if (t->format == RD_TRAJ_PDB) pdb_lines(buf[c], a0, a1, xyz, type);
        else xml_lines(buf[c], a0, a1, xyz, type);
    };
    if (nchunk == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int c = 0; c < nchunk; ++c) th.emplace_back(work, c);
        for (auto& x : th) x.join();
    }
    if (fwrite(head.data(), 1, head.size(), t->f) != head.size()) return RD_EINVAL;
    for (const std::string& b : buf)
        if (!b.empty() && fwrite(b.data(), 1, b.size(), t->f) != b.size()) return RD_EINVAL;
    if (fwrite(tail.data(), 1, tail.size(), t->f) != tail.size()) return RD_EINVAL;
    return RD_OK;
}

extern "C" int rd_traj_close(rd_traj* t)
{
    if (!t) return RD_EINVAL;
    int rc = RD_OK;
    if (t->format == RD_TRAJ_XML && fputs("</traj>\n", t->f) < 0) rc = RD_EINVAL;
    if (fclose(t->f) != 0) rc = RD_EINVAL;
    delete t;
    return rc;
}

extern "C" int rd_rdf_write(const char* path, int n_types, const char* const* names, int nbins, double range,
                            const double* g)
{
    if (!path || n_types < 1 || nbins < 1 || !g) return RD_EINVAL;
    FILE* f = fopen(path, "w");
    if (!f) return RD_EINVAL;
    bool ok = fprintf(f, "d") >= 0;
    for (int i = 0; ok && i < n_types; ++i)
        for (int j = 0; ok && j < n_types; ++j) {
            std::string a = names && names[i] ? names[i] : std::to_string(i);
            std::string b = names && names[j] ? names[j] : std::to_string(j);
            ok = fprintf(f, ",\tRDF(%s,%s)", a.c_str(), b.c_str()) >= 0;
        }
    ok = ok && fprintf(f, "\n") >= 0;
    for (int k = 0; ok && k < nbins; ++k) {
        ok = fprintf(f, "%8.5f", (double)k * range / (double)nbins) >= 0;
        for (int i = 0; ok && i < n_types; ++i)
            for (int j = 0; ok && j < n_types; ++j)
                ok = fprintf(f, ", %8.3f", g[((size_t)i * n_types + j) * nbins + k]) >= 0;
        ok = ok && fprintf(f, "\n") >= 0;
    }
    if (fclose(f) != 0) ok = false;
    return ok ? RD_OK : RD_EINVAL;
}
