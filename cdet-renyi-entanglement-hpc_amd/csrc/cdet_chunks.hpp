// cdet_chunks.hpp -- Green tables on disk in the reference's chunk-directory
// layout, streamed to and from HBM (included by cdet_kernels.hip).
//
// Reference: include/cdet/chunked_load_data.h -- paths and names :123-199,
// manifest :333-359, job plans :754-882, sample files :947-996, corner files
// :998-1033, diagonal lines :1035-1093, coefficient files :1095-1158, loading
// :1372-1459.  File contents are raw FP64 in host byte order, no headers.
//
// Transfers go through two pinned host buffers: while one buffer's copy runs
// on the context stream the other is filled from (or written to) the file.
// Host memory is therefore bounded by the staging size, not the table size.

#include <cerrno>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace {
namespace chunkio {

namespace fs = std::filesystem;

// Failure with a C-ABI status code; the text is the reference's where the
// reference has one.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void io_error(const std::string& m) { throw Error(CDET_ERR_IO, m); }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(CDET_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr const char* kKindName[4] = {"twisted_up", "twisted_down", "ordinary_up", "ordinary_down"};
constexpr const char* kBranchName[2] = {"upper", "lower"};
constexpr const char* kMaskName[6] = {"empty", "checkerboard", "first_row", "first_column", "origin_square",
                                      "left_columns"};

std::string lower_name(std::string s) {  // lower case, '-' -> '_'
    for (char& c : s) {
        c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
        if (c == '-') c = '_';
    }
    return s;
}

int kind_of(const std::string& raw) {
    const std::string n = lower_name(raw);
    for (int k = 0; k < 4; ++k)
        if (n == kKindName[k]) return k;
    io_error("Unknown chunked table kind: " + raw);
}

int branch_of(const std::string& raw) {
    const std::string n = lower_name(raw);
    if (n == "upper") return 0;
    if (n == "lower") return 1;
    io_error("Unknown chunked branch: " + raw);
}

// SUBREGION_MASK names and the reference's aliases (replica_free_greens.h:45-67).
int mask_of(const std::string& raw) {
    const std::string n = lower_name(raw);
    if (n == "empty" || n == "none" || n == "ordinary") return 0;
    if (n == "checkerboard") return 1;
    if (n == "first_row") return 2;
    if (n == "first_column") return 3;
    if (n == "origin_square" || n == "center_square" || n == "centered_square") return 4;
    if (n == "left_columns") return 5;
    io_error("Unknown SUBREGION_MASK: " + raw);
}

// "<begin>_<end>", both zero-padded to 10 digits (:129-134).
std::string range_token(uint64_t b, uint64_t e) {
    char buf[48];
    std::snprintf(buf, sizeof(buf), "%010llu_%010llu", static_cast<unsigned long long>(b),
                  static_cast<unsigned long long>(e));
    return buf;
}

fs::path sample_path(const fs::path& d, int kind, int br, uint64_t b, uint64_t e) {
    return d / "samples" / kKindName[kind] / kBranchName[br] / ("nodes_" + range_token(b, e) + ".bin");
}
fs::path coeff_path(const fs::path& d, int kind, int br, uint64_t b, uint64_t e) {
    return d / "coeffs" / kKindName[kind] / kBranchName[br] / ("sitepairs_" + range_token(b, e) + ".bin");
}
fs::path corner_path(const fs::path& d, int kind, int br) {
    return d / "corners" / kKindName[kind] / (std::string(kBranchName[br]) + ".bin");
}
fs::path diag_path(const fs::path& d, int kind, uint64_t b, uint64_t e) {
    return d / "diagonal_lines" / kKindName[kind] / ("sitepairs_" + range_token(b, e) + ".bin");
}

// Shortest decimal (%.{p}g, p = 1..17) that reads back to v.
std::string decimal(double v) {
    char buf[64];
    for (int p = 1; p <= 17; ++p) {
        std::snprintf(buf, sizeof(buf), "%.*g", p, v);
        if (std::strtod(buf, nullptr) == v) return buf;
    }
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    return buf;
}

struct Geometry {
    uint64_t sites, sitepairs, nodes;
};

Geometry geometry(const cdet_chunk_spec& s) {
    const uint64_t S = static_cast<uint64_t>(s.N1) * static_cast<uint64_t>(s.N2);
    return Geometry{S, 4 * S * S, static_cast<uint64_t>(s.cheb_nu) * static_cast<uint64_t>(s.cheb_nv)};
}

void check_spec(const cdet_chunk_spec& s) {
    if (s.N1 <= 0 || s.N2 <= 0 || s.cheb_nu <= 0 || s.cheb_nv <= 0)
        throw Error(CDET_ERR_ARGUMENT, "chunk spec: lattice and Chebyshev sizes must be positive");
    if (s.node_chunk_size <= 0 || s.sitepair_chunk_size <= 0)
        throw Error(CDET_ERR_ARGUMENT, "chunk spec: chunk sizes must be positive");
    if (s.mask_kind < 0 || s.mask_kind > 5) throw Error(CDET_ERR_ARGUMENT, "chunk spec: unknown mask kind");
}

// current_chunked_manifest_values (:333-359); std::map keeps the keys sorted
// as the reference's writer does.
std::map<std::string, std::string> manifest_of(const cdet_chunk_spec& s) {
    const Geometry g = geometry(s);
    return {
        {"layout_version", "1"},
        {"N1", std::to_string(s.N1)},
        {"N2", std::to_string(s.N2)},
        {"n_sites", std::to_string(g.sites)},
        {"replica_sitepair_count", std::to_string(g.sitepairs)},
        {"CHEB_NU", std::to_string(s.cheb_nu)},
        {"CHEB_NV", std::to_string(s.cheb_nv)},
        {"node_count", std::to_string(g.nodes)},
        {"node_chunk_size", std::to_string(s.node_chunk_size)},
        {"sitepair_chunk_size", std::to_string(s.sitepair_chunk_size)},
        {"BETA", decimal(s.beta)},
        {"MU", decimal(s.mu)},
        {"ALPHA_UP", decimal(s.alpha_up)},
        {"ALPHA_DOWN", decimal(s.alpha_down)},
        {"PBC", s.pbc ? "true" : "false"},
        {"SUBREGION_MASK", kMaskName[s.mask_kind]},
        {"SUBREGION_SQUARE_RADIUS", std::to_string(s.subregion_square_radius)},
        {"LOAD_DATA_REAL", "double"},
        {"VECTOR_REAL", "double"},
    };
}

// write_key_value_file (:383-398): temporary file, then rename.
void write_kv(const fs::path& path, const std::map<std::string, std::string>& kv) {
    fs::create_directories(path.parent_path());
    const fs::path tmp = path.string() + ".tmp";
    {
        std::ofstream out(tmp);
        if (!out.is_open()) io_error("Could not write " + tmp.string());
        for (const auto& [k, v] : kv) out << k << '=' << v << '\n';
        if (!out) io_error("Could not write " + tmp.string());
    }
    fs::rename(tmp, path);
}

// read_key_value_file (:361-381): '#' lines and lines without '=' skipped.
std::map<std::string, std::string> read_kv(const fs::path& path) {
    std::ifstream in(path);
    if (!in.is_open()) io_error("Could not open " + path.string());
    std::map<std::string, std::string> kv;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        const size_t pos = line.find('=');
        if (pos == std::string::npos) continue;
        kv[line.substr(0, pos)] = line.substr(pos + 1);
    }
    return kv;
}

struct Job {
    int index, kind, branch;
    uint64_t begin, end;
};

// node_jobs.tsv / transform_jobs.tsv (5 columns) and corner_jobs.tsv (3)
// (:455-538).
std::vector<Job> read_jobs(const fs::path& path, int columns, const char* what) {
    std::ifstream in(path);
    if (!in.is_open()) io_error("Could not open " + path.string());
    std::vector<Job> jobs;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::vector<std::string> tok;
        std::stringstream ss(line);
        std::string t;
        while (std::getline(ss, t, '\t')) tok.push_back(t);
        if (static_cast<int>(tok.size()) < columns) io_error(std::string("Malformed ") + what + " job line: " + line);
        Job j{};
        try {
            j.index = std::stoi(tok[0]);
            if (columns == 5) {
                j.begin = std::stoull(tok[3]);
                j.end = std::stoull(tok[4]);
            }
        } catch (const std::exception&) {
            io_error(std::string("Malformed ") + what + " job line: " + line);
        }
        j.kind = kind_of(tok[1]);
        j.branch = branch_of(tok[2]);
        jobs.push_back(j);
    }
    return jobs;
}

void write_jobs(const fs::path& path, const char* header, const std::vector<Job>& jobs, bool ranges) {
    std::ofstream out(path);
    if (!out.is_open()) io_error("Could not write " + path.string());
    out << header;
    for (const Job& j : jobs) {
        out << j.index << '\t' << kKindName[j.kind] << '\t' << kBranchName[j.branch];
        if (ranges) out << '\t' << j.begin << '\t' << j.end;
        out << '\n';
    }
    if (!out) io_error("Could not write " + path.string());
}

// Job plans of run_chunked_plan_for_real (:754-824): for every kind, then
// branch, consecutive ranges of `chunk`, indices from 1.
std::vector<Job> plan_ranges(uint64_t total, uint64_t chunk) {
    std::vector<Job> jobs;
    int index = 1;
    for (int k = 0; k < 4; ++k)
        for (int br = 0; br < 2; ++br)
            for (uint64_t b = 0; b < total; b += chunk) jobs.push_back(Job{index++, k, br, b, std::min(total, b + chunk)});
    return jobs;
}

void write_plan(const fs::path& dir, const cdet_chunk_spec& s) {
    check_spec(s);
    const Geometry g = geometry(s);
    fs::create_directories(dir);
    write_kv(dir / "manifest.txt", manifest_of(s));
    const std::vector<Job> nodes = plan_ranges(g.nodes, static_cast<uint64_t>(s.node_chunk_size));
    const std::vector<Job> trans = plan_ranges(g.sitepairs, static_cast<uint64_t>(s.sitepair_chunk_size));
    std::vector<Job> corners;
    int index = 1;
    for (int k = 0; k < 4; ++k)
        for (int br = 0; br < 2; ++br) corners.push_back(Job{index++, k, br, 0, 0});
    write_jobs(dir / "node_jobs.tsv", "#index\ttable\tbranch\tnode_begin\tnode_end\n", nodes, true);
    write_jobs(dir / "transform_jobs.tsv", "#index\ttable\tbranch\tsitepair_begin\tsitepair_end\n", trans, true);
    write_jobs(dir / "corner_jobs.tsv", "#index\ttable\tbranch\n", corners, false);
    // Resource summary (:826-881), FP64 sizes in GiB.
    const double gib = 1024.0 * 1024.0 * 1024.0;
    std::ofstream out(dir / "resource_plan.txt");
    if (!out.is_open()) io_error("Could not write " + (dir / "resource_plan.txt").string());
    out << "real_type=double\n";
    out << "node_count=" << g.nodes << '\n';
    out << "replica_sitepair_count=" << g.sitepairs << '\n';
    out << "node_chunk_size=" << s.node_chunk_size << '\n';
    out << "sitepair_chunk_size=" << s.sitepair_chunk_size << '\n';
    out << "node_jobs=" << nodes.size() << '\n';
    out << "transform_jobs=" << trans.size() << '\n';
    out << "corner_jobs=" << corners.size() << '\n';
    out << "sample_chunk_file_gib_estimate="
        << static_cast<double>(s.node_chunk_size) * static_cast<double>(g.sitepairs) * 8.0 / gib << '\n';
    out << "device_table_gib_per_kind="
        << 2.0 * static_cast<double>(g.sitepairs) * (static_cast<double>(g.nodes) + 1.0) * 8.0 / gib << '\n';
}

int64_t parse_int(const std::map<std::string, std::string>& kv, const char* key) {
    const auto it = kv.find(key);
    if (it == kv.end()) io_error(std::string("Chunk manifest is missing ") + key);
    try {
        size_t pos = 0;
        const long long v = std::stoll(it->second, &pos);
        if (pos != it->second.size()) throw std::invalid_argument(key);
        return v;
    } catch (const std::exception&) {
        io_error(std::string("Chunk manifest has a malformed ") + key + ": " + it->second);
    }
}

double parse_real(const std::map<std::string, std::string>& kv, const char* key) {
    const auto it = kv.find(key);
    if (it == kv.end()) io_error(std::string("Chunk manifest is missing ") + key);
    errno = 0;
    char* end = nullptr;
    const double v = std::strtod(it->second.c_str(), &end);
    if (end == it->second.c_str() || *end != '\0' || errno == ERANGE)
        io_error(std::string("Chunk manifest has a malformed ") + key + ": " + it->second);
    return v;
}

cdet_chunk_spec read_spec(const fs::path& dir) {
    const auto kv = read_kv(dir / "manifest.txt");
    cdet_chunk_spec s{};
    const auto lv = kv.find("layout_version");
    if (lv == kv.end() || lv->second != "1")
        io_error("Chunk manifest mismatch for layout_version: expected 1, found " +
                 (lv == kv.end() ? std::string("(missing)") : lv->second));
    for (const char* key : {"LOAD_DATA_REAL", "VECTOR_REAL"}) {
        const auto it = kv.find(key);
        if (it == kv.end() || lower_name(it->second) != "double")
            io_error(std::string("Chunk manifest mismatch for ") + key + ": expected double, found " +
                     (it == kv.end() ? std::string("(missing)") : it->second));
    }
    s.N1 = static_cast<int32_t>(parse_int(kv, "N1"));
    s.N2 = static_cast<int32_t>(parse_int(kv, "N2"));
    s.cheb_nu = static_cast<int32_t>(parse_int(kv, "CHEB_NU"));
    s.cheb_nv = static_cast<int32_t>(parse_int(kv, "CHEB_NV"));
    s.node_chunk_size = parse_int(kv, "node_chunk_size");
    s.sitepair_chunk_size = parse_int(kv, "sitepair_chunk_size");
    s.beta = parse_real(kv, "BETA");
    s.mu = parse_real(kv, "MU");
    s.alpha_up = parse_real(kv, "ALPHA_UP");
    s.alpha_down = parse_real(kv, "ALPHA_DOWN");
    const auto pbc = kv.find("PBC");
    if (pbc == kv.end()) io_error("Chunk manifest is missing PBC");
    const std::string pv = lower_name(pbc->second);
    s.pbc = (pv == "true" || pv == "1" || pv == "yes" || pv == "on") ? 1 : 0;
    const auto mask = kv.find("SUBREGION_MASK");
    if (mask == kv.end()) io_error("Chunk manifest is missing SUBREGION_MASK");
    s.mask_kind = mask_of(mask->second);
    s.subregion_square_radius = static_cast<int32_t>(parse_int(kv, "SUBREGION_SQUARE_RADIUS"));
    check_spec(s);
    // derived counts must agree with the sizes
    const Geometry g = geometry(s);
    const std::pair<const char*, uint64_t> derived[] = {
        {"n_sites", g.sites}, {"replica_sitepair_count", g.sitepairs}, {"node_count", g.nodes}};
    for (const auto& [key, want] : derived) {
        const int64_t got = parse_int(kv, key);
        if (static_cast<uint64_t>(got) != want)
            io_error(std::string("Chunk manifest mismatch for ") + key + ": expected " + std::to_string(want) +
                     ", found " + std::to_string(got));
    }
    return s;
}

template <typename T>
void expect_eq(const char* key, const T& want, const T& got) {
    if (!(want == got)) {
        std::ostringstream e;
        e << "Chunk manifest mismatch for " << key << ": expected " << want << ", found " << got;
        io_error(e.str());
    }
}

// ensure_manifest_matches_for_real (:884-907) against a context (and lattice).
cdet_chunk_spec check_against(const fs::path& dir, const cdet_params& p, const cdet_lattice* expect) {
    const cdet_chunk_spec s = read_spec(dir);
    expect_eq("N1", p.N1, s.N1);
    expect_eq("N2", p.N2, s.N2);
    expect_eq("CHEB_NU", p.cheb_nu, s.cheb_nu);
    expect_eq("CHEB_NV", p.cheb_nv, s.cheb_nv);
    expect_eq("BETA", decimal(p.beta), decimal(s.beta));
    expect_eq("MU", decimal(p.mu), decimal(s.mu));
    expect_eq("ALPHA_UP", decimal(p.alpha_up), decimal(s.alpha_up));
    expect_eq("ALPHA_DOWN", decimal(p.alpha_down), decimal(s.alpha_down));
    if (expect) {
        expect_eq("PBC", std::string(expect->pbc ? "true" : "false"), std::string(s.pbc ? "true" : "false"));
        if (expect->mask_kind < 0 || expect->mask_kind > 5) throw Error(CDET_ERR_ARGUMENT, "unknown mask kind");
        expect_eq("SUBREGION_MASK", std::string(kMaskName[expect->mask_kind]), std::string(kMaskName[s.mask_kind]));
        if (s.mask_kind == 4 || s.mask_kind == 5) expect_eq("SUBREGION_SQUARE_RADIUS", expect->mask_param, s.subregion_square_radius);
    }
    return s;
}

uint64_t size_of_file(const fs::path& path) {
    std::error_code ec;
    const uint64_t n = fs::file_size(path, ec);
    if (ec) io_error("Could not stat " + path.string() + ": " + ec.message());
    return n;
}

// require_file_size (:431-440).
void require_size(const fs::path& path, uint64_t expected) {
    const uint64_t got = size_of_file(path);
    if (got != expected) {
        std::ostringstream e;
        e << "Unexpected file size for " << path << ": expected " << expected << " bytes, got " << got;
        io_error(e.str());
    }
}

// Two pinned buffers with an event each for host <-> device streaming.
class Staging {
public:
    Staging(size_t bytes, hipStream_t s) : bytes_(bytes), s_(s) {
        for (int i = 0; i < 2; ++i) {
            hip_check(hipHostMalloc(&buf_[i], bytes_, hipHostMallocDefault), "hipHostMalloc");
            hip_check(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming), "hipEventCreate");
        }
    }
    ~Staging() {
        (void)hipStreamSynchronize(s_);
        for (int i = 0; i < 2; ++i) {
            if (ev_[i]) (void)hipEventDestroy(ev_[i]);
            if (buf_[i]) (void)hipHostFree(buf_[i]);
        }
    }
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;
    size_t bytes() const { return bytes_; }

    // Next free buffer: waits until the copy that last used it has finished.
    char* acquire() {
        slot_ ^= 1;
        if (pending_[slot_]) hip_check(hipEventSynchronize(ev_[slot_]), "hipEventSynchronize");
        pending_[slot_] = false;
        return static_cast<char*>(buf_[slot_]);
    }
    void h2d(void* dst, size_t n) {
        hip_check(hipMemcpyAsync(dst, buf_[slot_], n, hipMemcpyHostToDevice, s_), "hipMemcpyAsync");
        hip_check(hipEventRecord(ev_[slot_], s_), "hipEventRecord");
        pending_[slot_] = true;
    }
    // Copy n bytes device -> current buffer and wait for them.
    char* d2h(const void* src, size_t n) {
        char* h = acquire();
        hip_check(hipMemcpyAsync(h, src, n, hipMemcpyDeviceToHost, s_), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(s_), "hipStreamSynchronize");
        return h;
    }

private:
    size_t bytes_;
    hipStream_t s_;
    void* buf_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
    bool pending_[2] = {false, false};
    int slot_ = 1;
};

constexpr size_t kStagingBytes = size_t{64} << 20;

// Stream `count` doubles of a file (from byte `offset`) into device memory.
void file_to_device(Staging& st, const fs::path& path, uint64_t offset, double* dst, uint64_t count) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) io_error("Could not open " + path.string());
    in.seekg(static_cast<std::streamoff>(offset), std::ios::beg);
    const uint64_t per = st.bytes() / sizeof(double);
    for (uint64_t done = 0; done < count;) {
        const uint64_t n = std::min(per, count - done);
        char* h = st.acquire();
        in.read(h, static_cast<std::streamsize>(n * sizeof(double)));
        if (!in) io_error("Failed while reading chunked binary values from " + path.string());
        st.h2d(dst + done, n * sizeof(double));
        done += n;
    }
}

// Write `count` doubles of device memory to a file (temporary, then rename).
void device_to_file(Staging& st, const double* src, uint64_t count, const fs::path& path) {
    fs::create_directories(path.parent_path());
    const fs::path tmp = path.string() + ".tmp";
    {
        std::ofstream out(tmp, std::ios::binary);
        if (!out.is_open()) io_error("Could not write " + tmp.string());
        const uint64_t per = st.bytes() / sizeof(double);
        for (uint64_t done = 0; done < count;) {
            const uint64_t n = std::min(per, count - done);
            const char* h = st.d2h(src + done, n * sizeof(double));
            out.write(h, static_cast<std::streamsize>(n * sizeof(double)));
            if (!out) io_error("Failed while writing chunked binary values to " + tmp.string());
            done += n;
        }
    }
    fs::rename(tmp, path);
}

// Kinds that own storage after a load: with equal spin shifts the down kinds
// alias the up kinds (load_data.h:70-91).
bool aliased_down(const cdet_params& p, int kind) {
    return p.alpha_up == p.alpha_down && (kind == CDET_TABLE_TWISTED_DOWN || kind == CDET_TABLE_ORDINARY_DOWN);
}

// Zeroed device table for `kind` (the reference's destination vectors start at zero).
DevTable& fresh_table(cdet_ctx* ctx, int kind) {
    for (int k = 0; k < 4; ++k)
        if (k != kind && ctx->tab[k].alias == kind) {
            // a kind aliasing this one is re-pointed by the caller afterwards
            ctx->tab[k] = DevTable{};
        }
    if (alloc_table(ctx, kind)) throw Error(CDET_ERR_DEVICE, g_last_error);
    DevTable& t = ctx->tab[kind];
    invalidate_split(t);
    size_t nc, nk;
    table_sizes(ctx, nc, nk);
    hip_check(hipMemsetAsync(t.upper, 0, nc * sizeof(double), ctx->stream), "hipMemsetAsync");
    hip_check(hipMemsetAsync(t.lower, 0, nc * sizeof(double), ctx->stream), "hipMemsetAsync");
    hip_check(hipMemsetAsync(t.ucorner, 0, nk * sizeof(double), ctx->stream), "hipMemsetAsync");
    hip_check(hipMemsetAsync(t.lcorner, 0, nk * sizeof(double), ctx->stream), "hipMemsetAsync");
    return t;
}

void finish_load(cdet_ctx* ctx) {
    const cdet_params& p = ctx->p;
    if (p.alpha_up == p.alpha_down) {
        if (cdet_tables_alias(ctx, CDET_TABLE_TWISTED_DOWN, CDET_TABLE_TWISTED_UP) ||
            cdet_tables_alias(ctx, CDET_TABLE_ORDINARY_DOWN, CDET_TABLE_ORDINARY_UP))
            throw Error(CDET_ERR_TABLES, g_last_error);
    }
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    ctx->ordinary_layout = CDET_ORDINARY_CHUNKED;
    refresh_scalars(ctx);
}

void check_range(const Job& j, uint64_t total, const char* what) {
    if (!(j.begin < j.end && j.end <= total)) {
        std::ostringstream e;
        e << what << " job " << j.index << " has an out-of-range or empty interval [" << j.begin << ", " << j.end
          << ")";
        io_error(e.str());
    }
}

// load_chunked_data_into_resident_tables (:1372-1459), into HBM.
void load_tables(cdet_ctx* ctx, const fs::path& dir, const cdet_lattice* expect) {
    if (!fs::exists(dir / "READY"))
        io_error("Chunked load-data directory is missing READY: " + (dir / "READY").string());
    const cdet_chunk_spec spec = check_against(dir, ctx->p, expect);
    const Geometry g = geometry(spec);
    const std::vector<Job> jobs = read_jobs(dir / "transform_jobs.tsv", 5, "transform");
    for (const Job& j : jobs) check_range(j, g.sitepairs, "transform");
    Staging st(kStagingBytes, ctx->stream);
    for (int kind = 0; kind < 4; ++kind) {
        if (aliased_down(ctx->p, kind)) continue;
        DevTable& t = fresh_table(ctx, kind);
        for (const Job& j : jobs) {
            if (j.kind != kind) continue;
            const uint64_t n = (j.end - j.begin) * g.nodes;
            const fs::path path = coeff_path(dir, kind, j.branch, j.begin, j.end);
            require_size(path, n * sizeof(double));
            file_to_device(st, path, 0, (j.branch == 0 ? t.upper : t.lower) + j.begin * g.nodes, n);
        }
        for (int br = 0; br < 2; ++br) {
            const fs::path path = corner_path(dir, kind, br);
            require_size(path, g.sitepairs * sizeof(double));
            file_to_device(st, path, 0, br == 0 ? t.ucorner : t.lcorner, g.sitepairs);
        }
    }
    finish_load(ctx);
}

// Sample values of `count` nodes, node-major in `stage` ([j][sitepair]), to
// their [sitepair][node0 + j] places in a table.
__global__ __launch_bounds__(256) void k_place_nodes(const double* __restrict__ stage, int64_t count, int64_t sp,
                                                     int64_t node0, int64_t ncoef, double* __restrict__ dst) {
    const int64_t total = count * sp;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t j = e / sp;
        const int64_t f = e - j * sp;
        dst[f * ncoef + node0 + j] = stage[e];
    }
}

// The reverse: [sitepair][node0 + j] of a sample grid to node-major `stage`.
__global__ __launch_bounds__(256) void k_gather_nodes(const double* __restrict__ src, int64_t count, int64_t sp,
                                                      int64_t node0, int64_t ncoef, double* __restrict__ stage) {
    const int64_t total = count * sp;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t j = e / sp;
        const int64_t f = e - j * sp;
        stage[e] = src[f * ncoef + node0 + j];
    }
}

unsigned grid_for(int64_t n) { return static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 4096)); }

// Device buffer freed at scope exit (after the stream drains).
struct DevBuf {
    void* p = nullptr;
    hipStream_t s = nullptr;
    DevBuf(size_t bytes, hipStream_t st) : s(st) { hip_check(hipMalloc(&p, bytes), "hipMalloc"); }
    ~DevBuf() {
        if (p) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(p);
        }
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    double* d() const { return static_cast<double*>(p); }
};

// Whole nodes per staging piece for a node-major file.
uint64_t nodes_per_piece(uint64_t sitepairs) {
    return std::max<uint64_t>(1, kStagingBytes / (sitepairs * sizeof(double)));
}

// The transform stage (run_chunked_transform_sitepair_job_for_real,
// :1095-1158) for whole tables on the GPU: samples in, cosine transform in HBM.
void transform_samples(cdet_ctx* ctx, const fs::path& dir, const cdet_lattice* expect) {
    const cdet_chunk_spec spec = check_against(dir, ctx->p, expect);
    const Geometry g = geometry(spec);
    const std::vector<Job> jobs = read_jobs(dir / "node_jobs.tsv", 5, "node");
    for (const Job& j : jobs) check_range(j, g.nodes, "node");
    const uint64_t piece = nodes_per_piece(g.sitepairs);
    Staging st(std::max<size_t>(kStagingBytes, piece * g.sitepairs * sizeof(double)), ctx->stream);
    DevBuf stage(piece * g.sitepairs * sizeof(double), ctx->stream);
    double *d_nu = nullptr, *d_nv = nullptr, *d_cu = nullptr, *d_cv = nullptr;
    const int nrc = upload_node_tables(ctx->p, &d_nu, &d_nv, &d_cu, &d_cv);
    struct Free {
        double** p[4];
        ~Free() {
            for (double** x : p)
                if (*x) (void)hipFree(*x);
        }
    } free_nodes{{&d_nu, &d_nv, &d_cu, &d_cv}};
    if (nrc) throw Error(nrc, g_last_error);
    hipStream_t s = ctx->stream;
    for (int kind = 0; kind < 4; ++kind) {
        if (aliased_down(ctx->p, kind)) continue;
        DevTable& t = fresh_table(ctx, kind);
        for (int br = 0; br < 2; ++br) {
            double* dst = br == 0 ? t.upper : t.lower;
            for (const Job& j : jobs) {
                if (j.kind != kind || j.branch != br) continue;
                const fs::path path = sample_path(dir, kind, br, j.begin, j.end);
                require_size(path, (j.end - j.begin) * g.sitepairs * sizeof(double));
                for (uint64_t n0 = j.begin; n0 < j.end; n0 += piece) {
                    const uint64_t cnt = std::min(piece, j.end - n0);
                    file_to_device(st, path, (n0 - j.begin) * g.sitepairs * sizeof(double), stage.d(),
                                   cnt * g.sitepairs);
                    hipLaunchKernelGGL(k_place_nodes, dim3(grid_for(static_cast<int64_t>(cnt * g.sitepairs))),
                                       dim3(256), 0, s, stage.d(), static_cast<int64_t>(cnt),
                                       static_cast<int64_t>(g.sitepairs), static_cast<int64_t>(n0),
                                       static_cast<int64_t>(g.nodes), dst);
                    hip_check(hipGetLastError(), "k_place_nodes");
                }
            }
            const int rc = table_dct(ctx, d_cu, d_cv, dst, static_cast<int64_t>(g.sitepairs), s);
            if (rc) throw Error(rc, g_last_error);
            const fs::path cp = corner_path(dir, kind, br);
            require_size(cp, g.sitepairs * sizeof(double));
            file_to_device(st, cp, 0, br == 0 ? t.ucorner : t.lcorner, g.sitepairs);
        }
    }
    finish_load(ctx);
}

void write_ready(const fs::path& dir) {
    std::ofstream out(dir / "READY");
    if (!out.is_open()) io_error("Could not write " + (dir / "READY").string());
    out << "ready\n";
}

void check_spec_matches(const cdet_params& p, const cdet_chunk_spec& s) {
    check_spec(s);
    if (s.N1 != p.N1 || s.N2 != p.N2 || s.cheb_nu != p.cheb_nu || s.cheb_nv != p.cheb_nv || s.beta != p.beta ||
        s.mu != p.mu || s.alpha_up != p.alpha_up || s.alpha_down != p.alpha_down)
        throw Error(CDET_ERR_ARGUMENT, "chunk spec does not match the context parameters");
}

// Diagonal-line records of one lower-branch coefficient chunk
// (write_chunked_diagonal_line_file, :1035-1093): for every sitepair with
// equal replicas and equal sites, the row sums over v of its coefficients.
struct DiagLines {
    std::vector<uint64_t> index;
    std::vector<double> lines;  // nu per record
};

bool diagonal_lower_line(uint64_t f, int N1, int N2) {
    const int y2 = static_cast<int>(f % N2);
    f /= N2;
    const int x2 = static_cast<int>(f % N1);
    f /= N1;
    const int y1 = static_cast<int>(f % N2);
    f /= N2;
    const int x1 = static_cast<int>(f % N1);
    f /= N1;
    const int rb = static_cast<int>(f % 2);
    const int ra = static_cast<int>(f / 2);
    return ra == rb && x1 == x2 && y1 == y2;
}

void write_diag_file(const fs::path& path, const DiagLines& d, int nu) {
    fs::create_directories(path.parent_path());
    const fs::path tmp = path.string() + ".tmp";
    {
        std::ofstream out(tmp, std::ios::binary);
        if (!out.is_open()) io_error("Could not write " + tmp.string());
        const uint64_t n = d.index.size();
        out.write(reinterpret_cast<const char*>(&n), sizeof(n));
        for (size_t r = 0; r < d.index.size(); ++r) {
            out.write(reinterpret_cast<const char*>(&d.index[r]), sizeof(uint64_t));
            out.write(reinterpret_cast<const char*>(d.lines.data() + r * static_cast<size_t>(nu)),
                      static_cast<std::streamsize>(sizeof(double) * static_cast<size_t>(nu)));
        }
        if (!out) io_error("Failed while writing diagonal lines to " + tmp.string());
    }
    fs::rename(tmp, path);
}

// Device tables -> chunk directory.
void write_tables(cdet_ctx* ctx, const fs::path& dir, const cdet_chunk_spec& spec) {
    check_spec_matches(ctx->p, spec);
    if (!tables_ready(ctx)) throw Error(CDET_ERR_TABLES, "Resident Green-function tables have not been uploaded");
    if (ctx->ordinary_layout != CDET_ORDINARY_CHUNKED)
        throw Error(CDET_ERR_TABLES,
                    "ordinary tables are in the resident replica-(0,0) layout; build them with "
                    "ordinary_layout = CDET_ORDINARY_CHUNKED to write a chunk directory");
    write_plan(dir, spec);
    const Geometry g = geometry(spec);
    const std::vector<Job> jobs = read_jobs(dir / "transform_jobs.tsv", 5, "transform");
    const uint64_t sp_bytes = g.nodes * sizeof(double);
    const uint64_t per_piece = std::max<uint64_t>(1, kStagingBytes / sp_bytes);
    Staging st(std::max<size_t>(kStagingBytes, per_piece * sp_bytes), ctx->stream);
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    const int nu = spec.cheb_nu, nv = spec.cheb_nv;
    for (int kind = 0; kind < 4; ++kind) {
        const int own = ctx->tab[kind].alias >= 0 ? ctx->tab[kind].alias : kind;
        const DevTable& t = ctx->tab[own];
        for (const Job& j : jobs) {
            if (j.kind != kind) continue;
            const double* src = (j.branch == 0 ? t.upper : t.lower) + j.begin * g.nodes;
            const fs::path path = coeff_path(dir, kind, j.branch, j.begin, j.end);
            fs::create_directories(path.parent_path());
            const fs::path tmp = path.string() + ".tmp";
            DiagLines diag;
            {
                std::ofstream out(tmp, std::ios::binary);
                if (!out.is_open()) io_error("Could not write " + tmp.string());
                for (uint64_t f0 = j.begin; f0 < j.end; f0 += per_piece) {
                    const uint64_t cnt = std::min(per_piece, j.end - f0);
                    const double* h = reinterpret_cast<const double*>(
                        st.d2h(src + (f0 - j.begin) * g.nodes, cnt * sp_bytes));
                    out.write(reinterpret_cast<const char*>(h), static_cast<std::streamsize>(cnt * sp_bytes));
                    if (!out) io_error("Failed while writing chunked binary values to " + tmp.string());
                    if (j.branch != 1) continue;
                    for (uint64_t k = 0; k < cnt; ++k) {
                        if (!diagonal_lower_line(f0 + k, spec.N1, spec.N2)) continue;
                        diag.index.push_back(f0 + k);
                        const double* c = h + k * g.nodes;
                        for (int u = 0; u < nu; ++u) {
                            double sum = 0.0;
                            for (int v = 0; v < nv; ++v) sum += c[static_cast<size_t>(u) * nv + v];
                            diag.lines.push_back(sum);
                        }
                    }
                }
            }
            fs::rename(tmp, path);
            if (j.branch == 1) write_diag_file(diag_path(dir, kind, j.begin, j.end), diag, nu);
        }
        for (int br = 0; br < 2; ++br)
            device_to_file(st, br == 0 ? t.ucorner : t.lcorner, g.sitepairs, corner_path(dir, kind, br));
    }
    write_ready(dir);
}

// Sample and corner stages (:947-1033) by the GPU producer.
void write_samples(cdet_ctx* ctx, const fs::path& dir, const cdet_lattice* lat) {
    const cdet_chunk_spec spec = check_against(dir, ctx->p, lat);
    const Geometry g = geometry(spec);
    const std::vector<Job> jobs = read_jobs(dir / "node_jobs.tsv", 5, "node");
    for (const Job& j : jobs) check_range(j, g.nodes, "node");
    const uint64_t piece = nodes_per_piece(g.sitepairs);
    Staging st(std::max<size_t>(kStagingBytes, piece * g.sitepairs * sizeof(double)), ctx->stream);
    DevBuf stage(piece * g.sitepairs * sizeof(double), ctx->stream);
    DevBuf grid(g.sitepairs * g.nodes * sizeof(double), ctx->stream);
    DevBuf corner(g.sitepairs * sizeof(double), ctx->stream);
    hipStream_t s = ctx->stream;
    const cdet_params& p = ctx->p;
    for (int kind = 0; kind < 4; ++kind) {
        const bool twisted = kind < 2;
        const double mu_eff = p.mu - ((kind == 0 || kind == 2) ? p.alpha_up : p.alpha_down);
        Producer pr;
        int rc = producer_setup(ctx, mu_eff, twisted, lat, pr);
        if (rc) throw Error(rc, g_last_error);
        for (int br = 0; br < 2; ++br) {
            rc = producer_samples(ctx, pr, br == 0 ? 1 : 0, false, grid.d(), corner.d(), 0,
                                  static_cast<int>(g.nodes), lat->node_chunk, s);
            if (rc) throw Error(rc, g_last_error);
            rc = producer_status(pr, s);
            if (rc) throw Error(rc, g_last_error);
            for (const Job& j : jobs) {
                if (j.kind != kind || j.branch != br) continue;
                const fs::path path = sample_path(dir, kind, br, j.begin, j.end);
                fs::create_directories(path.parent_path());
                const fs::path tmp = path.string() + ".tmp";
                {
                    std::ofstream out(tmp, std::ios::binary);
                    if (!out.is_open()) io_error("Could not write " + tmp.string());
                    for (uint64_t n0 = j.begin; n0 < j.end; n0 += piece) {
                        const uint64_t cnt = std::min(piece, j.end - n0);
                        hipLaunchKernelGGL(k_gather_nodes, dim3(grid_for(static_cast<int64_t>(cnt * g.sitepairs))),
                                           dim3(256), 0, s, grid.d(), static_cast<int64_t>(cnt),
                                           static_cast<int64_t>(g.sitepairs), static_cast<int64_t>(n0),
                                           static_cast<int64_t>(g.nodes), stage.d());
                        hip_check(hipGetLastError(), "k_gather_nodes");
                        const char* h = st.d2h(stage.d(), cnt * g.sitepairs * sizeof(double));
                        out.write(h, static_cast<std::streamsize>(cnt * g.sitepairs * sizeof(double)));
                        if (!out) io_error("Failed while writing chunked binary values to " + tmp.string());
                    }
                }
                fs::rename(tmp, path);
            }
            device_to_file(st, corner.d(), g.sitepairs, corner_path(dir, kind, br));
        }
    }
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const Error& e) {
        return fail(e.code, e.what());
    } catch (const fs::filesystem_error& e) {
        return fail(CDET_ERR_IO, e.what());
    } catch (const std::exception& e) {
        return fail(CDET_ERR_IO, e.what());
    }
}

}  // namespace chunkio
}  // namespace

extern "C" int cdet_chunks_write_plan(const char* dir, const cdet_chunk_spec* spec) {
    if (!dir || !spec) return fail(CDET_ERR_ARGUMENT, "cdet_chunks_write_plan: bad argument");
    return chunkio::guarded([&] { chunkio::write_plan(dir, *spec); });
}

extern "C" int cdet_chunks_read_spec(const char* dir, cdet_chunk_spec* spec) {
    if (!dir || !spec) return fail(CDET_ERR_ARGUMENT, "cdet_chunks_read_spec: bad argument");
    return chunkio::guarded([&] { *spec = chunkio::read_spec(dir); });
}

extern "C" int cdet_chunks_load_tables(cdet_ctx* ctx, const char* dir, const cdet_lattice* expect) {
    if (!ctx || !dir) return fail(CDET_ERR_ARGUMENT, "cdet_chunks_load_tables: bad argument");
    return chunkio::guarded([&] { chunkio::load_tables(ctx, dir, expect); });
}

extern "C" int cdet_chunks_transform_samples(cdet_ctx* ctx, const char* dir, const cdet_lattice* expect) {
    if (!ctx || !dir) return fail(CDET_ERR_ARGUMENT, "cdet_chunks_transform_samples: bad argument");
    return chunkio::guarded([&] { chunkio::transform_samples(ctx, dir, expect); });
}

extern "C" int cdet_chunks_write_tables(cdet_ctx* ctx, const char* dir, const cdet_chunk_spec* spec) {
    if (!ctx || !dir || !spec) return fail(CDET_ERR_ARGUMENT, "cdet_chunks_write_tables: bad argument");
    return chunkio::guarded([&] { chunkio::write_tables(ctx, dir, *spec); });
}

extern "C" int cdet_chunks_write_samples(cdet_ctx* ctx, const char* dir, const cdet_lattice* lattice) {
    if (!ctx || !dir || !lattice) return fail(CDET_ERR_ARGUMENT, "cdet_chunks_write_samples: bad argument");
    return chunkio::guarded([&] { chunkio::write_samples(ctx, dir, lattice); });
}
