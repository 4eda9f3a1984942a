// er_host.cpp -- see er_host.h.  The loaders use the same stream-extraction
// rules as the reference so that malformed inputs behave the same way; the
// messages are the reference's, so a failing drop-in prints what the
// reference prints.
#include "er_host.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>
#include <thread>

namespace er {
namespace host {

namespace {

std::string join(const std::string &root, const std::string &rel) {
    if (root.empty() || root == ".") return rel;
    if (root.back() == '/') return root + rel;
    return root + "/" + rel;
}

std::string fmt(const char *f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char *f, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof(buf), f, ap);
    va_end(ap);
    return buf;
}

}  // namespace

int get_adsorbate_mw(const std::string &root, const std::string &adsorbate, double *mw,
                     std::string *msg) {
    std::ifstream massesfile(join(root, "data/masses.def").c_str());
    if (massesfile.fail()) {
        *msg = "Masses file data/masses.def not present!\n";
        return ST_IO;
    }
    double found = -1.0;
    std::string line, label;
    double mass = -1.0;
    std::getline(massesfile, line);  // header
    while (std::getline(massesfile, line)) {
        std::istringstream ls(line);
        ls >> label >> mass;
        if (label == adsorbate) {
            found = mass;
            break;
        }
    }
    if (found < 0.0) {
        *msg = fmt("No molecular weight for adsorbate %s in data/masses.def\n", adsorbate.c_str());
        return ST_MISSING;
    }
    *mw = found;
    return ST_OK;
}

int read_grid_settings(const std::string &root, GridParams *p, std::string *msg) {
    std::ifstream simfile(join(root, "simulation.input").c_str());
    if (simfile.fail()) {
        *msg = "Simulation.input failed to open!\n";
        return ST_IO;
    }
    p->forcefieldname = "None";
    p->verbose = false;
    p->grid_resolution = -1.0;
    p->gridoutputformat = "txt";
    p->r_cutoff_squared = 12.5 * 12.5;
    p->feynmanhibbs = false;
    p->energy_threshold = 0.0;
    p->T = -1;
    p->precision = 1;
    std::string word;
    while (simfile >> word) {
        if (word == "Temperature") simfile >> p->T;
        if (word == "VoidFractionEnergyThreshold") simfile >> p->energy_threshold;
        if (word == "GridResolution") simfile >> p->grid_resolution;
        if (word == "CutoffRadius") {
            simfile >> p->r_cutoff_squared;
            p->r_cutoff_squared = p->r_cutoff_squared * p->r_cutoff_squared;
        }
        if (word == "Forcefield") simfile >> p->forcefieldname;
        if (word == "GridOutputFormat") simfile >> p->gridoutputformat;
        if (word == "FeynmanHibbs") simfile >> p->feynmanhibbs;
        if (word == "verbosemode") simfile >> p->verbose;
        if (word == "AccessibilityGrid") simfile >> p->accessibility;  // extension
        if (word == "GridPrecision") {  // extension, ignored by the reference scanner
            std::string v;
            simfile >> v;
            if (v == "fp32" || v == "FP32") p->precision = 0;
            if (v == "fp64" || v == "FP64") p->precision = 1;
        }
    }
    if (p->feynmanhibbs & (p->T < 0.0)) {
        *msg = "Missing Temperature in simulation.input";
        return ST_MISSING;
    }
    if (p->grid_resolution < 0.0) {
        *msg = "Missing GridResolution in simulation.input";
        return ST_MISSING;
    }
    if (p->forcefieldname == "None") {
        *msg = "Missing Forcefield in simulation.input";
        return ST_MISSING;
    }
    return ST_OK;
}

int read_henry_settings(const std::string &root, HenryParams *p, std::string *msg) {
    std::ifstream simfile(join(root, "simulation.input").c_str());
    if (simfile.fail()) {
        *msg = "Simulation.input failed to open!\n";
        return ST_IO;
    }
    p->forcefieldname = "None";
    p->verbose = false;
    p->r_cutoff_squared = 12.5 * 12.5;
    p->T = -1;
    p->numinsertionsperA3 = -1;
    p->seed = 1234;
    std::string word;
    while (simfile >> word) {
        if (word == "Temperature") simfile >> p->T;
        if (word == "CutoffRadius") {
            simfile >> p->r_cutoff_squared;
            p->r_cutoff_squared = p->r_cutoff_squared * p->r_cutoff_squared;
        }
        if (word == "Forcefield") simfile >> p->forcefieldname;
        if (word == "NumberOfInsertionsPerA3") simfile >> p->numinsertionsperA3;
        if (word == "verbosemode") simfile >> p->verbose;
        if (word == "WidomSeed") simfile >> p->seed;  // extension
    }
    if (p->T < 0.0) {
        *msg = "Missing Temperature in simulation.input";
        return ST_MISSING;
    }
    if (p->numinsertionsperA3 < 0.0) {
        *msg = "Missing NumberOfInsertionsPerA3 in simulation.input";
        return ST_MISSING;
    }
    if (p->forcefieldname == "None") {
        *msg = "Missing Forcefield in simulation.input";
        return ST_MISSING;
    }
    return ST_OK;
}

int read_uc_file(const std::string &root, const std::string &framework,
                 const std::string &once_or_twice, int rep[3], std::string *msg) {
    std::ifstream ucfile(
        join(root, "data/uc_replications/" + framework + "_" + once_or_twice + ".uc").c_str());
    if (ucfile.fail()) {
        *msg = fmt("unit cell replication factor file not found in "
                   "data/uc_replications/$frameworkname_%s.uc\n",
                   once_or_twice.c_str());
        return ST_IO;
    }
    rep[0] = rep[1] = rep[2] = -1;
    for (int d = 0; d < 3; d++)
        if (!(ucfile >> rep[d])) {
            *msg = "Problem reading UC file";
            return ST_PARSE;
        }
    return ST_OK;
}

int load_forcefield(const std::string &root, const std::string &name, Forcefield *ff,
                    std::string *msg) {
    ff->name = name;
    const std::string rel = "data/forcefields/" + name + ".def";
    std::ifstream f(join(root, rel).c_str());
    if (f.fail()) {
        *msg = fmt("Forcefield file %s did not load.\n", rel.c_str());
        return ST_IO;
    }
    ff->numinteractions = 0;
    std::string line;
    std::getline(f, line);  // header
    while (std::getline(f, line)) ff->numinteractions++;
    f.clear();
    f.seekg(0, f.beg);
    std::getline(f, line);
    ff->epsilon.assign(ff->numinteractions, 0.0);
    ff->sigma.assign(ff->numinteractions, 0.0);
    ff->identity.assign(ff->numinteractions, std::string());
    for (int k = 0; k < ff->numinteractions; k++) {
        std::getline(f, line);
        std::istringstream ls(line);
        ls >> ff->identity[k] >> ff->epsilon[k] >> ff->sigma[k];
    }
    return ST_OK;
}

int load_framework(const std::string &root, const std::string &name, Framework *fw,
                   std::string *msg) {
    std::ifstream cssr(join(root, "data/structures/" + name + ".cssr").c_str());
    if (cssr.fail()) {
        *msg = fmt("CSSR file %s failed to import!", (name + ".cssr").c_str());
        return ST_IO;
    }
    fw->name = name;
    std::string line;
    std::getline(cssr, line);
    std::istringstream input(line);
    input >> fw->a >> fw->b >> fw->c;
    std::getline(cssr, line);
    input.str(line);
    input.clear();
    input >> fw->alpha >> fw->beta >> fw->gamma;
    fw->alpha = fw->alpha * M_PI / 180.0;
    fw->beta = fw->beta * M_PI / 180.0;
    fw->gamma = fw->gamma * M_PI / 180.0;
    std::getline(cssr, line);
    input.str(line);
    input.clear();
    input >> fw->noatoms;
    if (fw->noatoms < 0 || fw->noatoms > 50000000) {
        *msg = fmt("CSSR file %s has an invalid atom count!", (name + ".cssr").c_str());
        return ST_PARSE;
    }
    fw->x_f.assign(fw->noatoms, 0.0);
    fw->y_f.assign(fw->noatoms, 0.0);
    fw->z_f.assign(fw->noatoms, 0.0);
    fw->identity.assign(fw->noatoms, std::string());
    fw->mass.assign(fw->noatoms, 0.0);
    std::getline(cssr, line);  // skipped line
    for (int i = 0; i < fw->noatoms; i++) {
        std::getline(cssr, line);
        input.str(line);
        input.clear();
        int junk;
        input >> junk >> fw->identity[i] >> fw->x_f[i] >> fw->y_f[i] >> fw->z_f[i];
        // src/Framework.cc:58-60: C fmod keeps the sign of negative coordinates
        fw->x_f[i] = std::fmod(fw->x_f[i], 1.0);
        fw->y_f[i] = std::fmod(fw->y_f[i], 1.0);
        fw->z_f[i] = std::fmod(fw->z_f[i], 1.0);
    }
    const std::string massesfilename = "data/masses.def";
    std::ifstream massesfile(join(root, massesfilename).c_str());
    if (massesfile.fail()) {
        *msg = fmt("File %s not present.\n", massesfilename.c_str());
        return ST_IO;
    }
    std::getline(massesfile, line);
    int n_masses = 0;
    while (std::getline(massesfile, line)) n_masses++;
    massesfile.clear();
    massesfile.seekg(0, massesfile.beg);
    std::getline(massesfile, line);
    std::vector<double> masses(n_masses, 0.0);
    std::vector<std::string> identities(n_masses);
    for (int i = 0; i < n_masses; i++) {
        std::getline(massesfile, line);
        input.str(line);
        input.clear();
        input >> identities[i] >> masses[i];
    }
    for (int k = 0; k < fw->noatoms; k++) {
        bool found = false;
        for (int a = 0; a < n_masses; a++)
            if (fw->identity[k] == identities[a]) {
                found = true;
                fw->mass[k] = masses[a];
                break;
            }
        if (!found) {
            *msg = fmt("Atom %s not present in %s.\n", fw->identity[k].c_str(),
                       massesfilename.c_str());
            return ST_MISSING;
        }
    }
    // src/Framework.cc:113-119
    const double ca = std::cos(fw->alpha), cb = std::cos(fw->beta), cg = std::cos(fw->gamma);
    const double sg = std::sin(fw->gamma);
    const double v_unit_piped = std::sqrt(1.0 - std::pow(ca, 2.0) - std::pow(cb, 2) -
                                          std::pow(cg, 2) + 2 * ca * cb * cg);
    fw->volume_unitcell = v_unit_piped * fw->a * fw->b * fw->c;
    double mass_unitcell = 0.0;
    for (int k = 0; k < fw->noatoms; k++) mass_unitcell += fw->mass[k];
    fw->density = mass_unitcell / fw->volume_unitcell * 1660.53892;
    // src/Framework.cc:124-142
    double(*t)[3] = fw->t_matrix;
    t[0][0] = fw->a;
    t[0][1] = fw->b * cg;
    t[0][2] = fw->c * cb;
    t[1][0] = 0.0;
    t[1][1] = fw->b * sg;
    t[1][2] = fw->c * (ca - cb * cg) / sg;
    t[2][0] = 0.0;
    t[2][1] = 0.0;
    t[2][2] = fw->c * v_unit_piped / sg;
    double(*it)[3] = fw->inv_t_matrix;
    it[0][0] = 1.0 / fw->a;
    it[0][1] = -cg / fw->a / sg;
    it[0][2] = (ca * cg - cb) / (fw->a * v_unit_piped * sg);
    it[1][0] = 0.0;
    it[1][1] = 1.0 / fw->b / sg;
    it[1][2] = (cb * cg - ca) / (fw->b * v_unit_piped * sg);
    it[2][0] = 0.0;
    it[2][1] = 0.0;
    it[2][2] = sg / (fw->c * v_unit_piped);
    return ST_OK;
}

int grab_guest_params(const Forcefield &ff, const std::string &adsorbate, double *eps,
                      double *sig, std::string *msg) {
    for (int i = 0; i < ff.numinteractions; i++)
        if (ff.identity[i] == adsorbate) {
            *eps = ff.epsilon[i];
            *sig = ff.sigma[i];
            return ST_OK;
        }
    *msg = fmt("Could not find adsorbate %s in forcefield file %s", adsorbate.c_str(),
               ff.name.c_str());
    return ST_MISSING;
}

int load_particles(const Framework &fw, const Forcefield &ff, double eps_guest, double sig_guest,
                   std::vector<Particle> *out, std::string *msg) {
    out->assign(fw.noatoms, Particle());
    for (int i = 0; i < fw.noatoms; i++) {
        Particle &p = (*out)[i];
        p.x_f = fw.x_f[i];
        p.y_f = fw.y_f[i];
        p.z_f = fw.z_f[i];
        int k = 0;
        for (; k < ff.numinteractions; k++)
            if (ff.identity[k] == fw.identity[i]) break;
        if (k == ff.numinteractions) {
            *msg = fmt("Could not find framework atom %s in forcefield.\n", fw.identity[i].c_str());
            return ST_MISSING;
        }
        // Lorentz-Berthelot, src/load_fast_particle_f_array.h:35-36
        p.sig2 = std::pow((sig_guest + ff.sigma[k]) / 2.0, 2);
        p.eps = std::sqrt(eps_guest * ff.epsilon[k]);
    }
    return ST_OK;
}

void grid_dims(const Framework &fw, double resolution, int n[3]) {
    n[0] = static_cast<int>(std::ceil(fw.a / resolution));
    n[1] = static_cast<int>(std::ceil(fw.b / resolution));
    n[2] = static_cast<int>(std::ceil(fw.c / resolution));
}

void henry_counts(const Framework &fw, HenryParams *p) {
    p->numinsertions = static_cast<int>(p->numinsertionsperA3 * fw.volume_unitcell);
    p->numinsertionsperthread = p->numinsertions / (p->num_blocks * p->num_threads) + 1;
}

long long henry_total(const HenryParams &p) {
    return (long long)p.numinsertionsperthread * p.num_blocks * p.num_threads;
}

namespace {

struct BinHeader {
    char magic[8];
    uint32_t version, dtype;
    int32_t n[3];
    uint32_t reserved;
    double t[9];
    uint8_t pad[128 - 8 - 8 - 16 - 72];
};
static_assert(sizeof(BinHeader) == 128, "binary grid header is 128 bytes");
constexpr char kBinMagic[8] = {'E', 'R', 'G', 'R', 'I', 'D', '0', '1'};

bool slurp(const std::string &path, std::string *data) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    data->resize(sz > 0 ? (size_t)sz : 0);
    const size_t got = sz > 0 ? fread(&(*data)[0], 1, (size_t)sz, f) : 0;
    fclose(f);
    return got == data->size();
}

}  // namespace

int read_grid_txt(const std::string &path, int n[3], std::vector<double> *vals, int threads,
                  std::string *msg) {
    std::string data;
    if (!slurp(path, &data)) {
        *msg = fmt("Grid file %s could not be loaded...\n", path.c_str());  // src/gcmc.cc:297
        return ST_IO;
    }
    const size_t eol = data.find('\n');
    {
        std::istringstream head(data.substr(0, eol));
        head >> n[0] >> n[1] >> n[2];
        if (head.fail() || n[0] < 1 || n[1] < 1 || n[2] < 1 ||
            (long long)n[0] * n[1] * n[2] > (1LL << 36)) {
            *msg = fmt("Grid file %s has no valid N_x N_y N_z header", path.c_str());
            return ST_PARSE;
        }
    }
    const long long nlines = (long long)n[0] * n[1];
    std::vector<size_t> start;
    start.reserve((size_t)nlines + 1);
    size_t pos = eol == std::string::npos ? data.size() : eol + 1;
    while (pos < data.size() && (long long)start.size() < nlines) {
        start.push_back(pos);
        const void *nl = std::memchr(data.data() + pos, '\n', data.size() - pos);
        pos = nl ? (size_t)((const char *)nl - data.data()) + 1 : data.size();
    }
    if ((long long)start.size() < nlines) {
        *msg = fmt("Grid file %s has %lld of %lld z-pencil lines", path.c_str(),
                   (long long)start.size(), nlines);
        return ST_PARSE;
    }
    start.push_back(pos);
    vals->assign((size_t)nlines * n[2], 0.0);
    std::atomic<long long> bad{-1};
    auto work = [&](long long l0, long long l1) {
        for (long long l = l0; l < l1 && bad.load() < 0; l++) {
            const char *p = data.data() + start[l];
            const char *end = data.data() + start[l + 1];
            double *out = vals->data() + (size_t)l * n[2];
            for (int k = 0; k < n[2]; k++) {
                while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) p++;
                char *q = nullptr;
                const double v = std::strtod(p, &q);
                if (q == p || q > end || !std::isfinite(v)) {
                    long long expect = -1;
                    bad.compare_exchange_strong(expect, l);
                    return;
                }
                out[k] = v;
                p = q;
            }
        }
    };
    const int nt = std::max(1, std::min<int>(threads, (int)std::min<long long>(nlines, 64)));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
        pool.emplace_back(work, nlines * t / nt, nlines * (t + 1) / nt);
    for (auto &th : pool) th.join();
    if (bad.load() >= 0) {
        *msg = fmt("Grid file %s: line %lld does not hold %d values", path.c_str(),
                   bad.load() + 2, n[2]);
        return ST_PARSE;
    }
    return ST_OK;
}

int write_grid_bin(const std::string &path, const int n[3], const double t[9], const double *e64,
                   const float *e32, std::string *msg) {
    BinHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kBinMagic, 8);
    h.version = 1;
    h.dtype = e64 ? 1u : 0u;
    for (int d = 0; d < 3; d++) h.n[d] = n[d];
    std::memcpy(h.t, t, sizeof(h.t));
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
        *msg = fmt("could not write %s", path.c_str());
        return ST_IO;
    }
    const size_t np = (size_t)n[0] * n[1] * n[2];
    bool ok = fwrite(&h, sizeof(h), 1, f) == 1;
    ok = ok && (e64 ? fwrite(e64, sizeof(double), np, f) : fwrite(e32, sizeof(float), np, f)) == np;
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        *msg = fmt("could not write %s", path.c_str());
        return ST_IO;
    }
    return ST_OK;
}

int read_grid_bin(const std::string &path, int n[3], double t[9], int *dtype,
                  std::vector<char> *values, std::string *msg) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) {
        *msg = fmt("Grid file %s could not be loaded...\n", path.c_str());
        return ST_IO;
    }
    BinHeader h;
    const bool got = fread(&h, sizeof(h), 1, f) == 1;
    if (!got || std::memcmp(h.magic, kBinMagic, 8) != 0 || h.version != 1 || h.dtype > 1 ||
        h.n[0] < 1 || h.n[1] < 1 || h.n[2] < 1 ||
        (long long)h.n[0] * h.n[1] * h.n[2] > (1LL << 36)) {
        fclose(f);
        *msg = fmt("%s is not a binary grid file", path.c_str());
        return ST_PARSE;
    }
    for (int d = 0; d < 3; d++) n[d] = h.n[d];
    std::memcpy(t, h.t, sizeof(h.t));
    *dtype = (int)h.dtype;
    const size_t bytes = (size_t)n[0] * n[1] * n[2] * (h.dtype ? 8 : 4);
    values->resize(bytes);
    const bool ok = fread(values->data(), 1, bytes, f) == bytes;
    fclose(f);
    if (!ok) {
        *msg = fmt("binary grid file %s is truncated", path.c_str());
        return ST_PARSE;
    }
    return ST_OK;
}

int read_grid_file(const std::string &path, int n[3], std::vector<double> *vals, int threads,
                   std::string *msg) {
    char magic[8] = {0};
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) {
        *msg = fmt("Grid file %s could not be loaded...\n", path.c_str());
        return ST_IO;
    }
    const size_t got = fread(magic, 1, 8, f);
    fclose(f);
    if (got == 8 && std::memcmp(magic, kBinMagic, 8) == 0) {
        double t[9];
        int dtype = 0;
        std::vector<char> raw;
        int st = read_grid_bin(path, n, t, &dtype, &raw, msg);
        if (st) return st;
        const size_t np = (size_t)n[0] * n[1] * n[2];
        vals->resize(np);
        if (dtype) std::memcpy(vals->data(), raw.data(), np * sizeof(double));
        else
            for (size_t i = 0; i < np; i++) {
                float v;
                std::memcpy(&v, raw.data() + 4 * i, 4);
                (*vals)[i] = v;
            }
        return ST_OK;
    }
    return read_grid_txt(path, n, vals, threads, msg);
}

int write_accessibility_values(FILE *f, const double *acc, const int n[3]) {
    const long long syz = (long long)n[1] * n[2];
    for (int i = 0; i < n[0]; i++)
        for (int j = 0; j < n[1]; j++) {
            int count = 0;
            for (int k = 0; k < n[2]; k++) {
                fprintf(f, "%d ", int(acc[k + (long long)j * n[2] + (long long)i * syz]));
                if (++count == 6) {
                    fprintf(f, "\n");
                    count = 0;
                }
            }
            fprintf(f, "\n");
        }
    return ferror(f) ? ST_IO : ST_OK;
}

void flat_t(const Framework &fw, double t[9]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t[3 * i + j] = fw.t_matrix[i][j];
}

namespace {

void cross3(const double u[3], const double v[3], double w[3]) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

double dot3(const double u[3], const double v[3]) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; }

}  // namespace

int replication_factors(const double t[9], double cutoff, bool twice, int rep[3]) {
    const double rc = twice ? 2.0 * cutoff : cutoff;
    double v[3][3];  // cell vectors = columns of t
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < 3; i++) v[k][i] = t[3 * i + k];
    // normals of the faces spanned by the other two (original) vectors
    double nrm[3][3];
    cross3(v[1], v[2], nrm[0]);
    cross3(v[0], v[2], nrm[1]);
    cross3(v[0], v[1], nrm[2]);
    for (int k = 0; k < 3; k++) {
        rep[k] = 1;
        if (!(dot3(nrm[k], nrm[k]) > 0.0) || !(rc >= 0.0)) return ST_ARG;
    }
    for (int k = 0; k < 3; k++) {
        const double len = std::sqrt(dot3(nrm[k], nrm[k]));
        for (;;) {
            double c0[3];
            for (int i = 0; i < 3; i++) c0[i] = 0.5 * (v[0][i] + v[1][i] + v[2][i]);
            if (std::fabs(dot3(nrm[k], c0)) / len >= rc) break;
            if (rep[k] >= 64) return ST_ARG;
            rep[k]++;
            for (int i = 0; i < 3; i++) v[k][i] *= 2.0;
        }
    }
    return ST_OK;
}

int read_cssr_cell(const std::string &root, const std::string &name, double t[9],
                   std::string *msg) {
    std::ifstream cssr(join(root, "data/structures/" + name + ".cssr").c_str());
    if (cssr.fail()) {
        *msg = fmt("Crystal structure file %s not present in data/structures/",
                   (name + ".cssr").c_str());
        return ST_IO;
    }
    double a = 0, b = 0, c = 0, al = 0, be = 0, ga = 0;
    std::string line;
    std::getline(cssr, line);
    std::istringstream input(line);
    input >> a >> b >> c;
    std::getline(cssr, line);
    input.str(line);
    input.clear();
    input >> al >> be >> ga;
    if (input.fail()) {
        *msg = fmt("CSSR file %s has no cell", (name + ".cssr").c_str());
        return ST_PARSE;
    }
    al *= M_PI / 180.0;
    be *= M_PI / 180.0;
    ga *= M_PI / 180.0;
    const double ca = std::cos(al), cb = std::cos(be), cg = std::cos(ga), sg = std::sin(ga);
    const double vu = std::sqrt(1.0 - ca * ca - cb * cb - cg * cg + 2.0 * ca * cb * cg);
    const double m[9] = {a, b * cg, c * cb, 0.0, b * sg, c * (ca - cb * cg) / sg,
                         0.0, 0.0, c * vu / sg};
    std::memcpy(t, m, sizeof(m));
    return ST_OK;
}

int write_uc_file(const std::string &root, const std::string &name, bool twice,
                  const int rep[3], std::string *msg) {
    const std::string path =
        join(root, "data/uc_replications/" + name + (twice ? "_twice" : "_once") + ".uc");
    FILE *f = fopen(path.c_str(), "w");
    if (!f) {
        *msg = fmt("could not write %s", path.c_str());
        return ST_IO;
    }
    fprintf(f, "%d %d %d", rep[0], rep[1], rep[2]);
    fclose(f);
    return ST_OK;
}

namespace {

// Shared leading block of both .out headers
// (src/write_settings_to_outputfile.h:15-90 and :107-176).
void write_common_settings(FILE *f, const std::string &adsorbate, double mw, double eps_g,
                           double sig_g, double rc2, const int rep[3], const Framework &fw,
                           const Forcefield &ff, const Particle *atoms) {
    time_t now_t = time(0);
    struct tm *now = localtime(&now_t);
    fprintf(f, "date: %d-%d-%d\n\n", now->tm_year + 1900, now->tm_mon + 1, now->tm_mday);
    fprintf(f, "Adsorbate: %s\n", adsorbate.c_str());
    fprintf(f, "    Molecular weight = %f\n\n", mw);
    fprintf(f, "Structure: %s\n", fw.name.c_str());
    fprintf(f, "    Crystal density (kg/m3): %f\n", fw.density);
    fprintf(f, "    a = %f A; b = %f A; c = %f A\n    alpha = %f, beta = %f, gamma = %f\n", fw.a,
            fw.b, fw.c, fw.alpha / M_PI * 180.0, fw.beta / M_PI * 180.0,
            fw.gamma / M_PI * 180.0);
    fprintf(f, "    Atoms in framework: %d\n", fw.noatoms);
    fprintf(f, "    Atom    # in framework\n");
    for (int a = 0; a < ff.numinteractions; a++) {
        int count = 0;
        for (int k = 0; k < fw.noatoms; k++)
            if (fw.identity[k] == ff.identity[a]) count++;
        if (count != 0) fprintf(f, "    %s    %d\n", ff.identity[a].c_str(), count);
    }
    fprintf(f, "\n    Tranformation matrix:\n");
    for (int i = 0; i < 3; i++) {
        fprintf(f, "    ");
        for (int j = 0; j < 3; j++) fprintf(f, " %f ", fw.t_matrix[i][j]);
        fprintf(f, "\n");
    }
    fprintf(f, "\nForcefield: %s\n", ff.name.c_str());
    fprintf(f, "    %s epsilon (K) = %f, sigma (A) = %f\n", adsorbate.c_str(), eps_g, sig_g);
    fprintf(f, "    For LJ cutoff %f A, unit cell replication factors: %d %d %d\n", std::sqrt(rc2),
            rep[0], rep[1], rep[2]);
    fprintf(f, "    %s-XX interactions present:\n    XX    epsilon(K)    sigma(A)\n",
            adsorbate.c_str());
    std::vector<std::string> printed;
    for (int i = 0; i < fw.noatoms; i++) {
        bool found = false;
        for (const std::string &s : printed)
            if (s == fw.identity[i]) {
                found = true;
                break;
            }
        if (!found) {
            printed.push_back(fw.identity[i]);
            fprintf(f, "    %s    %f    %f\n", fw.identity[i].c_str(), atoms[i].eps,
                    std::sqrt(atoms[i].sig2));
        }
    }
}

}  // namespace

void write_grid_settings(FILE *f, const GridParams &p, const Framework &fw, const Forcefield &ff,
                         const Particle *atoms) {
    write_common_settings(f, p.adsorbate, p.adsorbateMW, p.epsilon_guest, p.sigma_guest,
                          p.r_cutoff_squared, p.rep, fw, ff, atoms);
    fprintf(f, "\nGrid: %d by %d by %d points. Total grid points = %lld\n", p.N_x, p.N_y, p.N_z,
            (long long)p.N_x * p.N_y * p.N_z);
    const double da = fw.a / (p.N_x - 1), db = fw.b / (p.N_y - 1), dc = fw.c / (p.N_z - 1);
    fprintf(f, "    Actual grid spacing: da = %f A, db = %f A, dc = %f A\n", da, db, dc);
}

void write_henry_settings(FILE *f, const HenryParams &p, const Framework &fw,
                          const Forcefield &ff, const Particle *atoms) {
    write_common_settings(f, p.adsorbate, p.adsorbateMW, p.epsilon_guest, p.sigma_guest,
                          p.r_cutoff_squared, p.rep, fw, ff, atoms);
    fprintf(f,
            "\nHenry info\n    # of blocks: %d\n    # of threads: %d\n    # of insertions per "
            "thread: %d\n",
            p.num_blocks, p.num_threads, p.numinsertionsperthread);
    fprintf(f, "    Specified # of insertions/ A3: %d\n", p.numinsertionsperA3);
    fprintf(f, "    Actual # of insertions/ A3: %f\n",
            p.numinsertionsperthread * p.num_blocks * p.num_threads / fw.volume_unitcell);
    fprintf(f, "    Total # of insertions: %lld\n", henry_total(p));
}

void write_txt_header(FILE *f, const int n[3]) {
    fprintf(f,
            "%d %d %d  = (N_x,N_y,N_z) grid points (grid is in fractional coords). Endpoints "
            "included.\n",
            n[0], n[1], n[2]);
}

void write_cube_header(FILE *f, const int n[3], const double t[3][3]) {
    fprintf(f, "\nThis is a grid file.\n");
    fprintf(f, "%d % 13.6lf % 13.6lf % 13.6lf\n", 0, 0.0, 0.0, 0.0);
    fprintf(f, "%d % 13.6lf % 13.6lf % 13.6lf\n", n[0], t[0][0] / (n[0] - 1), 0.0, 0.0);
    fprintf(f, "%d % 13.6lf % 13.6lf % 13.6lf\n", n[1], t[0][1] / (n[1] - 1),
            t[1][1] / (n[1] - 1), 0.0);
    fprintf(f, "%d % 13.6lf % 13.6lf % 13.6lf\n", n[2], t[0][2] / (n[2] - 1),
            t[1][2] / (n[2] - 1), t[2][2] / (n[2] - 1));
}

namespace {

void format_sheet(std::string *out, bool cube, const double *e64, const float *e32, const int n[3],
                  int i) {
    const int ny = n[1], nz = n[2];
    out->clear();
    out->reserve((size_t)ny * nz * 15 + ny * (cube ? nz / 6 + 2 : 1));
    char buf[64];
    const size_t base = (size_t)i * ny * nz;
    for (int j = 0; j < ny; j++) {
        int count = 0;
        for (int k = 0; k < nz; k++) {
            const size_t idx = base + (size_t)j * nz + k;
            double v = e64 ? e64[idx] : (double)e32[idx];
            if (cube) v = v * 8.314 / 1000;  // kJ/mol, src/writegrid.cu:248
            int len = snprintf(buf, sizeof(buf), "% 13.6E ", v);
            out->append(buf, (size_t)len);
            if (cube) {
                if (++count == 6) {
                    out->push_back('\n');
                    count = 0;
                }
            } else if (k == nz - 1) {
                out->push_back('\n');
            }
        }
        if (cube) out->push_back('\n');
    }
}

}  // namespace

int write_grid_values(FILE *f, const std::string &format, const double *e64, const float *e32,
                      const int n[3], int threads) {
    const bool cube = format == "cube";
    if (!cube && format != "txt") return ST_ARG;
    if (threads < 1) threads = 1;
    const int group = threads * 2;
    std::vector<std::string> bufs(group);
    for (int i0 = 0; i0 < n[0]; i0 += group) {
        const int cnt = std::min(group, n[0] - i0);
        std::vector<std::thread> pool;
        for (int w = 0; w < threads; w++)
            pool.emplace_back([&, w]() {
                for (int s = w; s < cnt; s += threads)
                    format_sheet(&bufs[s], cube, e64, e32, n, i0 + s);
            });
        for (auto &th : pool) th.join();
        for (int s = 0; s < cnt; s++)
            if (fwrite(bufs[s].data(), 1, bufs[s].size(), f) != bufs[s].size()) return ST_IO;
    }
    return ST_OK;
}

int setup_grid_run(const std::string &root, const std::string &structure,
                   const std::string &adsorbate, GridParams *p, RunSetup *r, std::string *msg) {
    // src/writegrid.cu:64-98, same order of checks
    p->frameworkname = structure;
    p->adsorbate = adsorbate;
    int st;
    if ((st = get_adsorbate_mw(root, adsorbate, &p->adsorbateMW, msg))) return st;
    if ((st = read_grid_settings(root, p, msg))) return st;
    if ((st = read_uc_file(root, structure, "once", p->rep, msg))) return st;
    if ((st = load_forcefield(root, p->forcefieldname, &r->ff, msg))) return st;
    if ((st = load_framework(root, structure, &r->fw, msg))) return st;
    p->N_framework_atoms = r->fw.noatoms;
    if ((st = grab_guest_params(r->ff, adsorbate, &p->epsilon_guest, &p->sigma_guest, msg)))
        return st;
    if ((st = load_particles(r->fw, r->ff, p->epsilon_guest, p->sigma_guest, &r->atoms, msg)))
        return st;
    int n[3];
    grid_dims(r->fw, p->grid_resolution, n);
    if (p->accessibility)  // src/writegrid.cu:104-108: a grid in fractional space
        for (int d = 0; d < 3; d++) n[d] = static_cast<int>(std::ceil(1.0 / p->grid_resolution));
    p->N_x = n[0];
    p->N_y = n[1];
    p->N_z = n[2];
    return ST_OK;
}

int setup_henry_run(const std::string &root, const std::string &structure,
                    const std::string &adsorbate, HenryParams *p, RunSetup *r, std::string *msg) {
    // src/henry.cu:50-95, same order of checks
    p->frameworkname = structure;
    p->adsorbate = adsorbate;
    int st;
    if ((st = get_adsorbate_mw(root, adsorbate, &p->adsorbateMW, msg))) return st;
    if ((st = read_henry_settings(root, p, msg))) return st;
    if ((st = read_uc_file(root, structure, "once", p->rep, msg))) return st;
    if ((st = load_forcefield(root, p->forcefieldname, &r->ff, msg))) return st;
    if ((st = load_framework(root, structure, &r->fw, msg))) return st;
    p->N_framework_atoms = r->fw.noatoms;
    henry_counts(r->fw, p);
    if ((st = grab_guest_params(r->ff, adsorbate, &p->epsilon_guest, &p->sigma_guest, msg)))
        return st;
    if ((st = load_particles(r->fw, r->ff, p->epsilon_guest, p->sigma_guest, &r->atoms, msg)))
        return st;
    return ST_OK;
}

}  // namespace host
}  // namespace er
