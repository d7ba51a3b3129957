// rd_xml.cpp -- reader for the reference's XML input files (the schema
// subset the GPU path reads), without Xerces-C.
//
// Follows XMLHandler/*.cpp of the reference for which elements and attributes
// matter; number lists are read by extracting every number, which also
// accepts the bracketed forms the reference's string helpers mis-read
// (AuxillaryFunctions.cpp:79-135: "[[1,1];[0,1]]" and "[1,2]").
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rdgpu.h"

namespace {

struct XNode {
    std::string name;
    std::map<std::string, std::string> attrs;
    std::string text;
    std::vector<XNode> kids;
    const XNode* child(const char* n) const
    {
        for (const XNode& k : kids)
            if (k.name == n) return &k;
        return nullptr;
    }
};

bool read_file(const char* path, std::string& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

// Minimal XML: elements, attributes ("..." or '...'), text, comments,
// processing instructions.  Entities are not expanded (the schema has none).
bool xml_parse(const std::string& s, XNode& root, std::string& err)
{
    std::vector<XNode*> stack;
    root = XNode();
    root.name = "#document";
    stack.push_back(&root);
    size_t i = 0, n = s.size();
    while (i < n) {
        if (s[i] != '<') {
            size_t j = s.find('<', i);
            if (j == std::string::npos) j = n;
            stack.back()->text.append(s, i, j - i);
            i = j;
            continue;
        }
        if (s.compare(i, 4, "<!--") == 0) {
            size_t j = s.find("-->", i + 4);
            if (j == std::string::npos) { err = "unterminated comment"; return false; }
            i = j + 3;
            continue;
        }
        if (s.compare(i, 2, "<?") == 0 || s.compare(i, 2, "<!") == 0) {
            size_t j = s.find('>', i);
            if (j == std::string::npos) { err = "unterminated declaration"; return false; }
            i = j + 1;
            continue;
        }
        if (s.compare(i, 2, "</") == 0) {
            size_t j = s.find('>', i);
            if (j == std::string::npos) { err = "unterminated end tag"; return false; }
            std::string nm = s.substr(i + 2, j - i - 2);
            while (!nm.empty() && isspace((unsigned char)nm.back())) nm.pop_back();
            if (stack.size() < 2 || stack.back()->name != nm) { err = "mismatched </" + nm + ">"; return false; }
            stack.pop_back();
            i = j + 1;
            continue;
        }
        // start tag
        size_t j = i + 1;
        while (j < n && !isspace((unsigned char)s[j]) && s[j] != '>' && s[j] != '/') ++j;
        XNode node;
        node.name = s.substr(i + 1, j - i - 1);
        bool selfclose = false;
        for (;;) {
            while (j < n && isspace((unsigned char)s[j])) ++j;
            if (j >= n) { err = "unterminated tag <" + node.name; return false; }
            if (s[j] == '>') { ++j; break; }
            if (s[j] == '/' && j + 1 < n && s[j + 1] == '>') { selfclose = true; j += 2; break; }
            size_t a = j;
            while (j < n && s[j] != '=' && !isspace((unsigned char)s[j]) && s[j] != '>') ++j;
            std::string an = s.substr(a, j - a);
            while (j < n && isspace((unsigned char)s[j])) ++j;
            if (j >= n || s[j] != '=') { err = "attribute without value in <" + node.name; return false; }
            ++j;
            while (j < n && isspace((unsigned char)s[j])) ++j;
            if (j >= n || (s[j] != '"' && s[j] != '\'')) { err = "unquoted attribute in <" + node.name; return false; }
            char q = s[j++];
            size_t b = s.find(q, j);
            if (b == std::string::npos) { err = "unterminated attribute in <" + node.name; return false; }
            node.attrs[an] = s.substr(j, b - j);
            j = b + 1;
        }
        stack.back()->kids.push_back(node);
        if (!selfclose) stack.push_back(&stack.back()->kids.back());
        i = j;
    }
    if (stack.size() != 1) { err = "unclosed <" + stack.back()->name + ">"; return false; }
    return true;
}

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a])) ++a;
    while (b > a && isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// every number in the string, in order
std::vector<double> numbers(const std::string& s)
{
    std::vector<double> v;
    const char* p = s.c_str();
    while (*p) {
        if (isdigit((unsigned char)*p) || ((*p == '-' || *p == '+' || *p == '.') && (isdigit((unsigned char)p[1]) || p[1] == '.'))) {
            char* e = nullptr;
            double x = strtod(p, &e);
            if (e == p) { ++p; continue; }
            v.push_back(x);
            p = e;
        } else {
            ++p;
        }
    }
    return v;
}

int subtype_id(const std::string& s)
{
    if (s == "attractive") return RD_SUB_ATTRACTIVE;
    if (s == "repulsive") return RD_SUB_REPULSIVE;
    if (s == "spring") return RD_SUB_SPRING;
    return 0;
}

}  // namespace

struct rd_model {
    rd_model_info info{};
    std::vector<std::string> type_names;
    std::vector<double> D, radius, radii;
    std::vector<float> xyz;
    std::vector<int> type, rid;
    std::vector<rd_pot1> pot1;
    std::vector<rd_pot2> pot2;
    std::vector<std::vector<uint8_t>> pot2mask;
    std::vector<int> bonds, bond_pot;    // bonded pairs of the group instances
    std::vector<rd_pot2> bpots;          // group potentials, templates in file order
    std::vector<double> bpot_r0;
};

static double numf(const XNode* n, double def)
{
    if (!n) return def;
    return atof(trim(n->text).c_str());
}

// every element named `name` anywhere below `n`
static void collect(const XNode& n, const char* name, std::vector<const XNode*>& out)
{
    for (const XNode& k : n.kids) {
        if (k.name == name) out.push_back(&k);
        collect(k, name, out);
    }
}

static int fail(char* err, int len, const std::string& m)
{
    if (err && len > 0) snprintf(err, (size_t)len, "%s", m.c_str());
    return RD_EINVAL;
}

extern "C" int rd_model_read_xml(const char* param_particles, const char* tplgy_potentials, const char* param_groups,
                                 const char* tplgy_coordinates, const char* tplgy_groups, const char* param_global,
                                 rd_model** out, char* err, int err_len)
{
    if (!out) return RD_EINVAL;
    *out = nullptr;
    rd_model* m = new rd_model();
    std::string txt, e;
    XNode doc;
    auto load = [&](const char* path, const char* what) -> bool {
        if (!path || !read_file(path, txt)) { fail(err, err_len, std::string("cannot read ") + what + ": " + (path ? path : "(null)")); return false; }
        if (!xml_parse(txt, doc, e)) { fail(err, err_len, std::string(what) + ": " + e); return false; }
        return true;
    };

    // --- param_global (GlobalParametersHandler.cpp:73-143) ---
    if (!load(param_global, "param_global")) { delete m; return RD_EINVAL; }
    {
        std::vector<const XNode*> v;
        auto get = [&](const char* name) -> const XNode* {
            v.clear();
            collect(doc, name, v);
            return v.empty() ? nullptr : v.back();
        };
        rd_model_info& I = m->info;
        I.frames = (int)numf(get("nSimulationSteps"), 0);
        I.dt_frame = numf(get("dt"), 0);
        I.dt_step = numf(get("dtOpenMM"), 0);
        if (get("dt_dynamic")) I.dt_step = numf(get("dt_dynamic"), I.dt_step);
        if (get("dt_dyn")) I.dt_step = numf(get("dt_dyn"), I.dt_step);
        I.device = (int)numf(get("cudaDeviceIndex"), 0);
        if (get("cudaDevice")) I.device = (int)numf(get("cudaDevice"), I.device);
        I.temperature = numf(get("T"), 0);
        I.boltzmann = numf(get("Kb"), 0);
        I.testmode = (int)numf(get("testmode"), 0) == 1;
        const XNode* lb = get("latticeBounds");
        std::vector<double> b = lb ? numbers(lb->text) : std::vector<double>();
        if (b.size() != 6) { delete m; return fail(err, err_len, "param_global: latticeBounds needs 6 numbers"); }
        for (int k = 0; k < 6; ++k) I.lattice_bounds[k] = b[k];
        if (!(I.dt_step > 0)) { delete m; return fail(err, err_len, "param_global: no BD step (dtOpenMM)"); }
        I.steps_per_frame = (int)std::lround(I.dt_frame / I.dt_step);
    }

    // --- param_particles (ParticleParametersHandler.cpp:48-162) ---
    if (!load(param_particles, "param_particles")) { delete m; return RD_EINVAL; }
    std::vector<std::vector<std::pair<std::string, double>>> cmaps;
    {
        std::vector<const XNode*> parts;
        collect(doc, "particle", parts);
        for (const XNode* p : parts) {
            const XNode* t = p->child("type");
            m->type_names.push_back(t ? trim(t->text) : "");
            m->D.push_back(numf(p->child("diffusionConstant"), 0.0));
            std::vector<std::pair<std::string, double>> cm;
            if (const XNode* cmap = p->child("collisionRadiusMap")) {
                for (const XNode& cr : cmap->kids) {
                    if (cr.name != "collisionRadius") continue;
                    const XNode* pt = cr.child("partnerType");
                    double r = numf(cr.child("radius"), 0.0);
                    std::string partners = pt ? trim(pt->text) : "default";
                    std::stringstream ss(partners);
                    std::string one;
                    while (std::getline(ss, one, ',')) cm.push_back({ trim(one), r });
                }
            }
            cmaps.push_back(cm);
        }
    }
    const int nt = (int)m->type_names.size();
    if (nt < 1 || nt > RD_MAX_TYPES) { delete m; return fail(err, err_len, "param_particles: 1.." + std::to_string(RD_MAX_TYPES) + " particle types"); }
    m->radius.assign(nt, 0.0);
    m->radii.assign((size_t)nt * nt, 0.0);
    // Simulation.cpp:208-229: "default"/"all" set the default radius, a partner
    // id sets one matrix entry (a partner type name is resolved as well)
    for (int i = 0; i < nt; ++i)
        for (auto& kv : cmaps[i]) {
            if (kv.first == "default" || kv.first == "all") { m->radius[i] = kv.second; continue; }
            int j = -1;
            for (int k = 0; k < nt; ++k)
                if (m->type_names[k] == kv.first) j = k;
            if (j < 0 && !kv.first.empty() && isdigit((unsigned char)kv.first[0])) j = atoi(kv.first.c_str());
            if (j >= 0 && j < nt) m->radii[(size_t)i * nt + j] = kv.second;
        }
    m->info.n_types = nt;

    // --- tplgy_potentials (PotentialParametersHandler.cpp:38-83, Simulation.cpp:231-252) ---
    if (!load(tplgy_potentials, "tplgy_potentials")) { delete m; return RD_EINVAL; }
    {
        std::vector<const XNode*> pots;
        collect(doc, "pot", pots);
        for (const XNode* p : pots) {
            auto A = [&](const char* k) -> std::string {
                auto it = p->attrs.find(k);
                return it == p->attrs.end() ? std::string() : trim(it->second);
            };
            auto vec3 = [&](const char* k, float* o) {
                std::vector<double> v = numbers(A(k));
                for (int d = 0; d < 3; ++d) o[d] = d < (int)v.size() ? (float)v[d] : 0.0f;
            };
            std::string type = A("type");
            int sub = subtype_id(A("subtype"));
            float k = (float)atof(A("forceConst").c_str());
            if (type == "DISK" || type == "CYLINDER" || type == "SPHERE" || type == "BOX" || type == "CUBE") {
                rd_pot1 q;
                memset(&q, 0, sizeof(q));
                q.type = type == "DISK" ? RD_POT1_DISK : type == "CYLINDER" ? RD_POT1_CYLINDER
                       : type == "SPHERE" ? RD_POT1_SPHERE : RD_POT1_BOX;
                q.subtype = sub;
                q.force_const = k;
                if (q.type == RD_POT1_BOX) { vec3("origin", q.origin); vec3("extension", q.extension); }
                else { vec3("center", q.origin); vec3("normal", q.normal); }
                q.radius = (float)atof(A("radius").c_str());
                q.height = (float)atof(A("height").c_str());
                std::string ids = A("affectedParticleTypeIds");
                if (ids == "all") q.type_mask = (1u << nt) - 1u;
                else
                    for (double t : numbers(ids))
                        if (t >= 0 && t < nt) q.type_mask |= 1u << (int)t;
                m->pot1.push_back(q);
            } else if (type == "HARMONIC" || type == "WEAK_INTERACTION_HARMONIC" ||
                       type == "WEAK_INTERACTION_PIECEWISE_HARMONIC") {
                rd_pot2 q;
                memset(&q, 0, sizeof(q));
                q.type = type == "HARMONIC" ? RD_POT2_HARMONIC
                       : type == "WEAK_INTERACTION_HARMONIC" ? RD_POT2_WEAK_HARMONIC : RD_POT2_PIECEWISE_WELL;
                q.subtype = sub;
                q.force_const = k;
                q.length = (float)atof(A("length").c_str());
                q.depth = (float)atof(A("depth").c_str());
                std::vector<uint8_t> mask((size_t)nt * nt, 0);
                std::string pairs = A("affectedParticleTypeIdPairs");
                if (pairs == "all") std::fill(mask.begin(), mask.end(), 1);
                else {
                    std::vector<double> v = numbers(pairs);
                    for (size_t t = 0; t + 1 < v.size(); t += 2) {
                        int a = (int)v[t], b = (int)v[t + 1];
                        if (a >= 0 && a < nt && b >= 0 && b < nt) mask[(size_t)a * nt + b] = mask[(size_t)b * nt + a] = 1;
                    }
                }
                if ((int)m->pot2.size() >= RD_MAX_POT2) { delete m; return fail(err, err_len, "too many order-two potentials"); }
                m->pot2.push_back(q);
                m->pot2mask.push_back(mask);
            } else {
                delete m;
                return fail(err, err_len, "tplgy_potentials: unknown potential type " + type);
            }
        }
        m->info.n_pot1 = (int)m->pot1.size();
        m->info.n_pot2 = (int)m->pot2.size();
    }

    // --- tplgy_coordinates (CoordReader.cpp:16-65): <p id type c="x,y,z"/> ---
    if (!load(tplgy_coordinates, "tplgy_coordinates")) { delete m; return RD_EINVAL; }
    {
        std::vector<const XNode*> ps;
        collect(doc, "p", ps);
        for (const XNode* p : ps) {
            auto it = p->attrs.find("c");
            std::vector<double> c = it == p->attrs.end() ? std::vector<double>() : numbers(it->second);
            if (c.size() != 3) { delete m; return fail(err, err_len, "tplgy_coordinates: bad c attribute"); }
            int t = atoi(p->attrs.count("type") ? p->attrs.at("type").c_str() : "0");
            if (t < 0 || t >= nt) { delete m; return fail(err, err_len, "tplgy_coordinates: unknown type " + std::to_string(t)); }
            m->xyz.push_back((float)c[0]); m->xyz.push_back((float)c[1]); m->xyz.push_back((float)c[2]);
            m->type.push_back(t);
            m->rid.push_back(atoi(p->attrs.count("id") ? p->attrs.at("id").c_str() : "0"));
        }
        m->info.n_particles = (int)m->type.size();
    }

    // --- param_groups (GroupParametersHandler.cpp, Simulation.cpp:254-264):
    //     one template per <particleGroup>, its <pot> entries with
    //     affectedInternalIdPairs; tplgy_groups (GroupReader.cpp:13-40):
    //     <g type="template" internalAndParticleId="[[internal,particleId];...]"/>.
    //     Every affected internal pair of every instance becomes a bond with
    //     that template's potential (the reference GPU code bonds only the
    //     first two members, always with potential 0).  Missing subtype:
    //     spring, missing forceConst: 10 (P2_Harmonic.java:47-49 defaults). ---
    struct Tmpl {
        int first = 0;                                  // index of its first potential in bpots
        std::vector<std::vector<std::pair<int, int>>> pairs;   // per potential: internal id pairs
    };
    std::vector<Tmpl> tmpls;
    // a missing group file means no groups, as GroupReader.cpp:17-19
    if (param_groups && param_groups[0] && read_file(param_groups, txt)) {
        if (!xml_parse(txt, doc, e)) { delete m; return fail(err, err_len, "param_groups: " + e); }
        std::vector<const XNode*> gs;
        collect(doc, "particleGroup", gs);
        for (const XNode* g : gs) {
            Tmpl T;
            T.first = (int)m->bpots.size();
            std::vector<const XNode*> pots;
            collect(*g, "pot", pots);
            for (const XNode* p : pots) {
                auto A = [&](const char* k) -> std::string {
                    auto it = p->attrs.find(k);
                    return it == p->attrs.end() ? std::string() : trim(it->second);
                };
                std::string type = A("type");
                rd_pot2 q;
                memset(&q, 0, sizeof(q));
                if (type == "HARMONIC") q.type = RD_POT2_HARMONIC;
                else if (type == "WEAK_INTERACTION_HARMONIC") q.type = RD_POT2_WEAK_HARMONIC;
                else if (type == "WEAK_INTERACTION_PIECEWISE_HARMONIC") q.type = RD_POT2_PIECEWISE_WELL;
                else continue;                              // not an order-two form (as Simulation.cpp:257-261)
                std::string sub = A("subtype");
                q.subtype = sub.empty() ? RD_SUB_SPRING : subtype_id(sub);
                std::string k = A("forceConst");
                q.force_const = k.empty() ? 10.0f : (float)atof(k.c_str());
                q.length = (float)atof(A("length").c_str());
                q.depth = (float)atof(A("depth").c_str());
                std::string r0 = A("r0");
                std::vector<double> v = numbers(A("affectedInternalIdPairs"));
                std::vector<std::pair<int, int>> pr;
                for (size_t t = 0; t + 1 < v.size(); t += 2) pr.emplace_back((int)v[t], (int)v[t + 1]);
                if ((int)m->bpots.size() >= RD_MAX_BOND_POTS) { delete m; return fail(err, err_len, "param_groups: too many group potentials"); }
                m->bpots.push_back(q);
                m->bpot_r0.push_back(r0.empty() || r0 == "null" ? -1.0 : atof(r0.c_str()));
                T.pairs.push_back(pr);
            }
            tmpls.push_back(T);
        }
    }
    int ngroups = 0;
    if (tplgy_groups && tplgy_groups[0] && read_file(tplgy_groups, txt)) {
        if (!xml_parse(txt, doc, e)) { delete m; return fail(err, err_len, "tplgy_groups: " + e); }
        std::map<int, int> index_of;                   // ReaDDy particle id -> index
        for (size_t i = 0; i < m->rid.size(); ++i) index_of[m->rid[i]] = (int)i;
        std::vector<const XNode*> gs;
        collect(doc, "g", gs);
        for (const XNode* g : gs) {
            ++ngroups;
            int t = atoi(g->attrs.count("type") ? g->attrs.at("type").c_str() : "-1");
            if (t < 0 || t >= (int)tmpls.size()) { delete m; return fail(err, err_len, "tplgy_groups: unknown group type " + std::to_string(t)); }
            std::vector<double> v = numbers(g->attrs.count("internalAndParticleId") ? g->attrs.at("internalAndParticleId") : "");
            std::map<int, int> member;                 // internal id -> particle index
            for (size_t k = 0; k + 1 < v.size(); k += 2) {
                auto it = index_of.find((int)v[k + 1]);
                if (it == index_of.end()) { delete m; return fail(err, err_len, "tplgy_groups: unknown particle id " + std::to_string((int)v[k + 1])); }
                member[(int)v[k]] = it->second;
            }
            const Tmpl& T = tmpls[t];
            for (size_t p = 0; p < T.pairs.size(); ++p)
                for (const auto& pr : T.pairs[p]) {
                    auto a = member.find(pr.first), c = member.find(pr.second);
                    if (a == member.end() || c == member.end()) { delete m; return fail(err, err_len, "tplgy_groups: a group lacks a member its potentials name"); }
                    m->bonds.push_back(a->second);
                    m->bonds.push_back(c->second);
                    m->bond_pot.push_back(T.first + (int)p);
                }
        }
    }
    m->info.n_groups = ngroups;
    m->info.n_bonds = (int)m->bond_pot.size();
    m->info.n_bond_pots = (int)m->bpots.size();
    *out = m;
    return RD_OK;
}

extern "C" void rd_model_free(rd_model* m) { delete m; }

extern "C" int rd_model_get_info(const rd_model* m, rd_model_info* info)
{
    if (!m || !info) return RD_EINVAL;
    *info = m->info;
    return RD_OK;
}

extern "C" int rd_model_types(const rd_model* m, double* D, double* radius, double* radii)
{
    if (!m) return RD_EINVAL;
    int nt = m->info.n_types;
    if (D) memcpy(D, m->D.data(), sizeof(double) * nt);
    if (radius) memcpy(radius, m->radius.data(), sizeof(double) * nt);
    if (radii) memcpy(radii, m->radii.data(), sizeof(double) * nt * nt);
    return RD_OK;
}

extern "C" int rd_model_particles(const rd_model* m, float* xyz, int* type, int* rid)
{
    if (!m) return RD_EINVAL;
    size_t n = m->type.size();
    if (xyz) memcpy(xyz, m->xyz.data(), sizeof(float) * 3 * n);
    if (type) memcpy(type, m->type.data(), sizeof(int) * n);
    if (rid) memcpy(rid, m->rid.data(), sizeof(int) * n);
    return RD_OK;
}

extern "C" int rd_model_pot1(const rd_model* m, int i, rd_pot1* p)
{
    if (!m || !p || i < 0 || i >= (int)m->pot1.size()) return RD_EINVAL;
    *p = m->pot1[i];
    return RD_OK;
}

extern "C" int rd_model_pot2(const rd_model* m, int i, rd_pot2* p, uint8_t* mask)
{
    if (!m || !p || i < 0 || i >= (int)m->pot2.size()) return RD_EINVAL;
    *p = m->pot2[i];
    if (mask) memcpy(mask, m->pot2mask[i].data(), m->pot2mask[i].size());
    return RD_OK;
}

extern "C" int rd_model_bonds(const rd_model* m, int* bonds, int* pot_index, rd_pot2* pots, double* r0)
{
    if (!m) return RD_EINVAL;
    if (bonds && !m->bonds.empty()) memcpy(bonds, m->bonds.data(), sizeof(int) * m->bonds.size());
    if (pot_index && !m->bond_pot.empty()) memcpy(pot_index, m->bond_pot.data(), sizeof(int) * m->bond_pot.size());
    if (pots && !m->bpots.empty()) memcpy(pots, m->bpots.data(), sizeof(rd_pot2) * m->bpots.size());
    if (r0 && !m->bpot_r0.empty()) memcpy(r0, m->bpot_r0.data(), sizeof(double) * m->bpot_r0.size());
    return RD_OK;
}

extern "C" int rd_create_from_model(const rd_model* m, double max_cutoff, uint64_t seed, rd_system** out)
{
    if (!m || !out) return RD_EINVAL;
    *out = nullptr;
    const rd_model_info& I = m->info;
    rd_params prm;
    memset(&prm, 0, sizeof(prm));
    for (int k = 0; k < 6; ++k) prm.lattice_bounds[k] = I.lattice_bounds[k];
    prm.dt = I.dt_step;
    prm.boltzmann = I.boltzmann;
    prm.temperature = I.temperature;
    prm.max_cutoff = max_cutoff;
    prm.cell_size = 0.0;
    prm.seed = seed;
    prm.device = I.device;
    rd_system* h = nullptr;
    int rc = rd_create(&prm, &h);
    if (rc) return rc;
    int nt = I.n_types;
    rc = rd_set_particle_types(h, nt, m->D.data(), m->radius.data(), m->radii.data());
    for (size_t k = 0; rc == RD_OK && k < m->pot1.size(); ++k) rc = rd_add_order_one_potential(h, &m->pot1[k]);
    for (size_t k = 0; rc == RD_OK && k < m->pot2.size(); ++k)
        rc = rd_add_order_two_potential(h, &m->pot2[k], m->pot2mask[k].data());
    if (rc == RD_OK) rc = rd_set_particles(h, I.n_particles, m->xyz.data(), m->type.data());
    if (rc == RD_OK && !m->bond_pot.empty())
        rc = rd_set_bonds(h, (int)m->bond_pot.size(), m->bonds.data(), m->bond_pot.data(), (int)m->bpots.size(),
                          m->bpots.data(), m->bpot_r0.data());
    if (rc == RD_OK) rc = rd_initialize(h);
    if (rc) {
        fprintf(stderr, "rd_create_from_model: %s\n", rd_last_error(h));
        rd_destroy(h);
        return rc;
    }
    *out = h;
    return RD_OK;
}
