// Simulation data files of the reference (include/tfs_io.h): writers that
// produce the bytes std::ofstream produces for the reference's records,
// readers of those records, the state -> source-position reconstruction of
// the flow-field program, and that program's save loop over tfs_flow.
//
// Host code only (built with the host compiler); the GPU is reached through
// the tfs_flow C-ABI.
#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tfs_flow.h"
#include "tfs_io.h"

namespace tfs {
int fail(int code, const std::string& msg);
}  // namespace tfs

using tfs::fail;

namespace {

// ---- formatting ---------------------------------------------------------
// std::scientific << std::setprecision(10) formats as printf's "%.10e" and
// the default float format with precision 6 as "%g"; std::to_chars with an
// explicit precision is specified to give printf's output for the same
// conversion.

void put_sci10(std::string& out, double v) {
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific, 10);
    out.append(buf, r.ptr);
}

void put_general6(std::string& out, double v) {
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::general, 6);
    out.append(buf, r.ptr);
}

void put_int(std::string& out, long long v) {
    char buf[32];
    auto r = std::to_chars(buf, buf + sizeof buf, v);
    out.append(buf, r.ptr);
}

int write_file(const char* path, const std::string& text, bool append, const char* who) {
    std::FILE* f = std::fopen(path, append ? "ab" : "wb");
    if (!f) return fail(TFS_ERR_IO, std::string(who) + ": cannot open " + path + ": " + std::strerror(errno));
    const size_t n = std::fwrite(text.data(), 1, text.size(), f);
    const int closed = std::fclose(f);
    if (n != text.size() || closed != 0)
        return fail(TFS_ERR_IO, std::string(who) + ": cannot write " + path);
    return TFS_OK;
}

// ---- parsing ------------------------------------------------------------
inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

inline const char* skip_space(const char* p, const char* e) {
    while (p < e && is_space(*p)) ++p;
    return p;
}

inline const char* token_end(const char* p, const char* e) {
    while (p < e && !is_space(*p)) ++p;
    return p;
}

// One number token as fscanf("%lf") reads it from the reference's files
// (decimal or exponent form, nan and inf, an optional sign).
bool parse_double(const char* p, const char* e, double* out) {
    if (p < e && *p == '+') ++p;
    auto r = std::from_chars(p, e, *out, std::chars_format::general);
    if (r.ec == std::errc::result_out_of_range && r.ptr == e && e - p < 64) {
        // Beyond the double range: fscanf gives +-HUGE_VAL or an underflowed
        // value there, as strtod does.
        char buf[64];
        std::memcpy(buf, p, size_t(e - p));
        buf[e - p] = 0;
        *out = std::strtod(buf, nullptr);
        return true;
    }
    return r.ec == std::errc() && r.ptr == e;
}

bool parse_int(const char* p, const char* e, long long* out) {
    if (p < e && *p == '+') ++p;
    auto r = std::from_chars(p, e, *out);
    return r.ec == std::errc() && r.ptr == e;
}

bool integer_token(const char* p, const char* e) {
    long long v;
    return parse_int(p, e, &v);
}

// Reads a whole line (without its newline) into `line`; false at the end of
// the file.
bool read_line(std::FILE* f, std::string& line) {
    line.clear();
    char buf[1 << 16];
    bool any = false;
    while (std::fgets(buf, sizeof buf, f)) {
        any = true;
        const size_t n = std::strlen(buf);
        if (n > 0 && buf[n - 1] == '\n') {
            line.append(buf, n - 1);
            return true;
        }
        line.append(buf, n);
    }
    return any;
}

bool blank(const std::string& s) {
    for (char c : s)
        if (!is_space(c)) return false;
    return true;
}

int64_t count_tokens(const std::string& s) {
    const char* p = s.data();
    const char* e = p + s.size();
    int64_t n = 0;
    for (p = skip_space(p, e); p < e; p = skip_space(token_end(p, e), e)) ++n;
    return n;
}

}  // namespace

struct tfs_io_reader {
    std::FILE* f = nullptr;
    std::string path;
    std::string line;       // the next record when `pending`
    bool pending = false;
    int64_t record = 0;     // records consumed (for messages)
};

namespace {

// Makes the next non-blank line the pending record; false at the end.
bool fill(tfs_io_reader* r) {
    if (r->pending) return true;
    while (read_line(r->f, r->line)) {
        if (!blank(r->line)) {
            r->pending = true;
            return true;
        }
    }
    return false;
}

// Unit-quaternion rotation matrix (quaternion.cpp:327-370), row-major.
void rot_mat(const double q[4], double R[9]) {
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    R[0] = 1.0 - 2.0 * y * y - 2.0 * z * z;
    R[4] = 1.0 - 2.0 * x * x - 2.0 * z * z;
    R[8] = 1.0 - 2.0 * x * x - 2.0 * y * y;
    R[1] = 2.0 * x * y - 2.0 * w * z;
    R[3] = 2.0 * x * y + 2.0 * w * z;
    R[2] = 2.0 * x * z + 2.0 * w * y;
    R[6] = 2.0 * x * z - 2.0 * w * y;
    R[5] = 2.0 * y * z - 2.0 * w * x;
    R[7] = 2.0 * y * z + 2.0 * w * x;
}

// First column of R(q): the segment tangent (quaternion.cpp:285-291).
void tangent(const double q[4], double t[3]) {
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    t[0] = 1.0 - 2.0 * (y * y + z * z);
    t[1] = 2.0 * (x * y + w * z);
    t[2] = 2.0 * (x * z - w * y);
}

void rotate_add(const double R[9], const double X[3], const double* v, double* out) {
    for (int i = 0; i < 3; ++i) out[i] = X[i] + (R[3 * i] * v[0] + R[3 * i + 1] * v[1] + R[3 * i + 2] * v[2]);
}

std::string dir_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

}  // namespace

extern "C" {

int tfs_io_write_par(const char* path, int32_t count, const tfs_io_param* params) {
    if (!path || count < 0 || (count > 0 && !params))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_write_par: null path or parameters");
    std::string text;
    for (int32_t i = 0; i < count; ++i) {
        const tfs_io_param& p = params[i];
        if (memchr(p.name, 0, sizeof p.name) == nullptr || p.name[0] == 0)
            return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_write_par: parameter name empty or unterminated");
        if (p.is_integer) {
            if (!(std::fabs(p.value) < 9.0e18) || p.value != std::floor(p.value))
                return fail(TFS_ERR_INVALID_ARGUMENT,
                            std::string("tfs_io_write_par: ") + p.name + " is not an integer");
            put_int(text, (long long)p.value);
        } else {
            put_general6(text, p.value);
        }
        text += " %% ";
        text += p.name;
        text += '\n';
    }
    return write_file(path, text, false, "tfs_io_write_par");
}

int tfs_io_read_par(const char* path, int32_t capacity, tfs_io_param* params, int32_t* count) {
    if (!path || !count || capacity < 0 || (capacity > 0 && !params))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_read_par: null argument");
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(TFS_ERR_IO, std::string("tfs_io_read_par: cannot open ") + path);
    std::string line;
    int32_t n = 0;
    int rc = TFS_OK;
    while (read_line(f, line)) {
        if (blank(line)) continue;
        const char* p = line.data();
        const char* e = p + line.size();
        const char* t0 = skip_space(p, e);
        const char* t1 = token_end(t0, e);
        const char* m0 = skip_space(t1, e);
        const char* m1 = token_end(m0, e);
        const char* n0 = skip_space(m1, e);
        const char* n1 = token_end(n0, e);
        double v;
        if (!parse_double(t0, t1, &v) || std::string(m0, m1) != "%%" || n0 == n1 ||
            skip_space(n1, e) != e || size_t(n1 - n0) >= sizeof(params->name)) {
            rc = fail(TFS_ERR_FORMAT, std::string("tfs_io_read_par: ") + path + ": line " +
                                          std::to_string(n + 1) + " is not 'value %% NAME'");
            break;
        }
        if (n < capacity) {
            tfs_io_param& q = params[n];
            std::memset(q.name, 0, sizeof q.name);
            std::memcpy(q.name, n0, size_t(n1 - n0));
            q.value = v;
            q.is_integer = integer_token(t0, t1) ? 1 : 0;
        }
        ++n;
    }
    std::fclose(f);
    *count = n;
    return rc;
}

int tfs_io_write_record(const char* path, int32_t flags, int64_t step, int64_t count, const double* values) {
    if (!path || count < 0 || (count > 0 && !values))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_write_record: null path or values");
    if (flags & ~(TFS_IO_APPEND | TFS_IO_WITH_STEP | TFS_IO_DEFAULT_FLOAT))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_write_record: unknown flags");
    std::string text;
    text.reserve(size_t(count) * 18 + 32);
    if (flags & TFS_IO_WITH_STEP) {
        put_int(text, (long long)step);
        text += ' ';
    }
    const bool general = (flags & TFS_IO_DEFAULT_FLOAT) != 0;
    for (int64_t i = 0; i < count; ++i) {
        if (general)
            put_general6(text, values[i]);
        else
            put_sci10(text, values[i]);
        text += ' ';
    }
    text += '\n';
    return write_file(path, text, (flags & TFS_IO_APPEND) != 0, "tfs_io_write_record");
}

int tfs_io_open(const char* path, tfs_io_reader** out) {
    if (!path || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_open: null argument");
    *out = nullptr;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(TFS_ERR_IO, std::string("tfs_io_open: cannot open ") + path);
    tfs_io_reader* r = new tfs_io_reader();
    r->f = f;
    r->path = path;
    *out = r;
    return TFS_OK;
}

int tfs_io_close(tfs_io_reader* r) {
    if (!r) return TFS_OK;
    if (r->f) std::fclose(r->f);
    delete r;
    return TFS_OK;
}

int tfs_io_rewind(tfs_io_reader* r) {
    if (!r) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_rewind: null reader");
    std::rewind(r->f);
    r->pending = false;
    r->record = 0;
    return TFS_OK;
}

int tfs_io_peek_count(tfs_io_reader* r, int64_t* count) {
    if (!r || !count) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_peek_count: null argument");
    if (!fill(r)) return TFS_IO_END;
    *count = count_tokens(r->line);
    return TFS_OK;
}

int tfs_io_read_record(tfs_io_reader* r, int32_t with_step, int64_t count, int64_t* step, double* values) {
    if (!r || count < 0 || (count > 0 && !values))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_read_record: null argument");
    if (!fill(r)) return TFS_IO_END;
    r->pending = false;
    const int64_t rec = r->record++;
    auto bad = [&](const std::string& why) {
        return fail(TFS_ERR_FORMAT, "tfs_io_read_record: " + r->path + ": record " + std::to_string(rec) + ": " + why);
    };
    const char* p = r->line.data();
    const char* e = p + r->line.size();
    p = skip_space(p, e);
    if (with_step) {
        const char* t = token_end(p, e);
        long long s;
        if (!parse_int(p, t, &s)) return bad("the step is not an integer");
        if (step) *step = s;
        p = skip_space(t, e);
    }
    for (int64_t i = 0; i < count; ++i) {
        if (p == e) return bad("holds " + std::to_string(i) + " values, expected " + std::to_string(count));
        const char* t = token_end(p, e);
        if (!parse_double(p, t, &values[i]))
            return bad("value " + std::to_string(i) + " '" + std::string(p, t) + "' is not a number");
        p = skip_space(t, e);
    }
    if (p != e)
        return bad("holds more than the expected " + std::to_string(count) + " values");
    return TFS_OK;
}

int tfs_io_count_records(const char* path, int64_t* count) {
    if (!path || !count) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_count_records: null argument");
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(TFS_ERR_IO, std::string("tfs_io_count_records: cannot open ") + path);
    std::string line;
    int64_t n = 0;
    while (read_line(f, line))
        if (!blank(line)) ++n;
    std::fclose(f);
    *count = n;
    return TFS_OK;
}

int tfs_io_sources_from_states(int32_t num_bodies, int32_t num_fil, int32_t num_seg, int32_t num_blobs,
                               double dl, int32_t quaternion_based, const double* body_states,
                               const double* seg_states, const double* fil_refs, const double* blob_refs,
                               double* x_segs, double* x_blobs, double* x_bodies) {
    if (num_bodies < 0 || num_fil < 0 || num_seg < 0 || num_blobs < 0)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_sources_from_states: negative size");
    const int64_t ns = int64_t(num_fil) * num_seg;
    if (!body_states || (ns > 0 && (!seg_states || !x_segs || (quaternion_based && !fil_refs))) ||
        (num_blobs > 0 && (!blob_refs || !x_blobs)))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_sources_from_states: null array");
    for (int32_t b = 0; b < num_bodies; ++b) {
        const double* X = body_states + 7 * int64_t(b);
        double R[9];
        rot_mat(X + 3, R);
        if (x_bodies)
            for (int i = 0; i < 3; ++i) x_bodies[3 * int64_t(b) + i] = X[i];
        for (int32_t f = 0; f < num_fil; ++f) {
            const int64_t s0 = int64_t(b) * ns + int64_t(f) * num_seg;
            double* xs = x_segs + 3 * s0;
            if (!quaternion_based) {
                std::memcpy(xs, seg_states + 3 * s0, sizeof(double) * 3 * size_t(num_seg));
                continue;
            }
            const double* qs = seg_states + 4 * s0;
            double pos[3], t1[3], t2[3];
            rotate_add(R, X, fil_refs + 3 * int64_t(f), pos);
            tangent(qs, t1);
            for (int i = 0; i < 3; ++i) xs[i] = pos[i];
            for (int32_t s = 1; s < num_seg; ++s) {
                tangent(qs + 4 * s, t2);
                for (int i = 0; i < 3; ++i) {
                    pos[i] += 0.5 * dl * (t1[i] + t2[i]);
                    xs[3 * s + i] = pos[i];
                    t1[i] = t2[i];
                }
            }
        }
        for (int32_t k = 0; k < num_blobs; ++k)
            rotate_add(R, X, blob_refs + 3 * int64_t(k), x_blobs + 3 * (int64_t(b) * num_blobs + k));
    }
    return TFS_OK;
}

int tfs_io_flow_run(const char* input_path, int32_t precision, const double* x_points, int64_t* saves) {
    if (!input_path) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_flow_run: null input path");
    if (saves) *saves = 0;
    // Parameter file: "KEY value" pairs in a fixed order (flow_field_main.cpp:13-20, 112-151).
    std::FILE* pf = std::fopen(input_path, "rb");
    if (!pf) return fail(TFS_ERR_IO, std::string("tfs_io_flow_run: cannot open ") + input_path);
    std::vector<std::string> tok;
    {
        std::string line;
        while (read_line(pf, line)) {
            const char* p = line.data();
            const char* e = p + line.size();
            for (p = skip_space(p, e); p < e;) {
                const char* t = token_end(p, e);
                tok.emplace_back(p, t);
                p = skip_space(t, e);
            }
        }
        std::fclose(pf);
    }
    if (tok.size() < 14) return fail(TFS_ERR_FORMAT, "tfs_io_flow_run: the parameter file needs 7 'KEY value' pairs");
    long long iv[6];
    for (int i = 0; i < 6; ++i)
        if (!parse_int(tok[3 + 2 * i].data(), tok[3 + 2 * i].data() + tok[3 + 2 * i].size(), &iv[i]))
            return fail(TFS_ERR_FORMAT, "tfs_io_flow_run: " + tok[2 + 2 * i] + " is not an integer");
    const int64_t nflow = iv[0];
    const bool quaternion_based = iv[1] != 0, periodic = iv[2] != 0;
    const int64_t cycles = periodic ? iv[3] : 1;
    const bool remove_net_force = iv[4] != 0, advance = iv[5] != 0;
    if (nflow < 1) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_io_flow_run: NFLOW must be positive");
    std::string name = tok[1];
    if (name.empty() || name[0] != '/') name = dir_of(input_path) + name;

    // Run parameters, by name (the reference reads them by position from the
    // same lines, flow_field_main.cpp:24-60).
    std::vector<tfs_io_param> par(64);
    int32_t npar = 0;
    int rc = tfs_io_read_par((name + ".par").c_str(), (int32_t)par.size(), par.data(), &npar);
    if (rc) return rc;
    auto get = [&](const char* key, double* v) {
        for (int32_t i = 0; i < std::min<int32_t>(npar, (int32_t)par.size()); ++i)
            if (std::strcmp(par[i].name, key) == 0) {
                *v = par[i].value;
                return true;
            }
        return false;
    };
    double NFIL, NSEG, NBOD, NBLOB, MU, RSEG, TOTAL, DT, PLOT;
    if (!get("NFIL", &NFIL) || !get("NSEG", &NSEG) || !get("NSWIM", &NBOD) || !get("NBLOB", &NBLOB) ||
        !get("MU", &MU) || !get("RSEG", &RSEG) || !get("TOTAL_TIME_STEPS", &TOTAL) || !get("DT", &DT) ||
        !get("PLOT_FREQUENCY_IN_STEPS", &PLOT))
        return fail(TFS_ERR_FORMAT, "tfs_io_flow_run: " + name + ".par lacks a parameter the run needs");
    const int nfil = (int)NFIL, nseg = (int)NSEG, nbod = (int)NBOD, nblob = (int)NBLOB;
    const int64_t plot = (int64_t)PLOT;
    if (nfil < 0 || nseg < 1 || nbod < 1 || nblob < 0 || plot < 1 || TOTAL < 1)
        return fail(TFS_ERR_FORMAT, "tfs_io_flow_run: " + name + ".par sizes out of range");
    const int64_t max_saves = 1 + ((int64_t)TOTAL - 1) / plot;
    const double dt_save = DT * double(plot);
    const double dl = 2.2 * RSEG;  // flow_field_evaluator.cu:410
    const int64_t ns = int64_t(nbod) * nfil * nseg, nb = int64_t(nbod) * nblob;
    const int seg_vals = quaternion_based ? 4 : 3;

    // Readers.
    struct Readers {
        tfs_io_reader *body = nullptr, *seg = nullptr, *force = nullptr, *blob = nullptr;
        ~Readers() {
            tfs_io_close(body);
            tfs_io_close(seg);
            tfs_io_close(force);
            tfs_io_close(blob);
        }
    } rd;
    if ((rc = tfs_io_open((name + "_body_states.dat").c_str(), &rd.body))) return rc;
    if ((rc = tfs_io_open((name + "_seg_states.dat").c_str(), &rd.seg))) return rc;
    if ((rc = tfs_io_open((name + "_seg_forces.dat").c_str(), &rd.force))) return rc;
    if (nblob > 0 && (rc = tfs_io_open((name + "_blob_forces.dat").c_str(), &rd.blob))) return rc;
    std::vector<double> fil_refs(3 * size_t(nfil)), blob_refs(3 * size_t(nblob));
    {
        tfs_io_reader* r = nullptr;
        if ((rc = tfs_io_open((name + "_fil_references.dat").c_str(), &r))) return rc;
        rc = tfs_io_read_record(r, 0, 3LL * nfil, nullptr, fil_refs.data());
        tfs_io_close(r);
        if (rc == TFS_IO_END) return fail(TFS_ERR_FORMAT, "tfs_io_flow_run: empty filament reference file");
        if (rc) return rc;
        if (nblob > 0) {
            if ((rc = tfs_io_open((name + "_blob_references.dat").c_str(), &r))) return rc;
            rc = tfs_io_read_record(r, 0, 3LL * nblob, nullptr, blob_refs.data());
            tfs_io_close(r);
            if (rc == TFS_IO_END) return fail(TFS_ERR_FORMAT, "tfs_io_flow_run: empty blob reference file");
            if (rc) return rc;
        }
    }

    // Evaluator (flow_field_evaluator.cu:265-337).
    tfs_flow_desc fd{};
    fd.num_bodies = nbod;
    fd.segs_per_body = nfil * nseg;
    fd.blobs_per_body = nblob;
    fd.mode = TFS_FLOW_STOKESLET_ROTLET;
    fd.precision = precision;
    fd.remove_net_force = remove_net_force ? 1 : 0;
    fd.seg_radius = RSEG;
    fd.blob_radius = RSEG;
    fd.mu = MU;
    fd.memory = TFS_MEMORY_HOST;
    fd.stream = nullptr;
    tfs_flow* flow = nullptr;
    if ((rc = tfs_flow_create(&fd, &flow))) return rc;
    struct FlowGuard {
        tfs_flow* h;
        ~FlowGuard() { tfs_flow_destroy(h); }
    } guard{flow};

    std::vector<double> xf(3 * size_t(nflow)), vf(3 * size_t(nflow), 0.0);
    if (x_points) {
        std::memcpy(xf.data(), x_points, xf.size() * sizeof(double));
    } else {
        for (int64_t n = 0; n < nflow; ++n) {  // flow_field_evaluator.cu:387-389
            const double row = std::floor(double(n) / 551.0);
            xf[3 * n] = 0.0;
            xf[3 * n + 1] = -2.2032e+03 + 8.0 * (double(n) - 551.0 * row);
            xf[3 * n + 2] = 8.0 * (1.0 + row);
        }
    }
    std::vector<double> body(7 * size_t(nbod)), seg(size_t(seg_vals) * ns), fseg(6 * size_t(ns)),
        fblob(3 * size_t(nb)), xs(3 * size_t(ns)), xb(3 * size_t(nb)), xbod(3 * size_t(nbod));
    const std::string loc_path = name + "_flow_locations.dat", vel_path = name + "_flow_velocities.dat";
    int64_t done = 0;
    for (int64_t cycle = 0; cycle < cycles; ++cycle) {
        bool ended = false;
        for (int64_t k = 0; k < max_saves; ++k) {
            int64_t step = 0;
            rc = tfs_io_read_record(rd.body, 1, 7LL * nbod, &step, body.data());
            if (rc == TFS_OK) rc = tfs_io_read_record(rd.seg, 1, seg_vals * ns, &step, seg.data());
            if (rc == TFS_OK) rc = tfs_io_read_record(rd.force, 1, 6 * ns, &step, fseg.data());
            if (rc == TFS_OK && nblob > 0) rc = tfs_io_read_record(rd.blob, 1, 3 * nb, &step, fblob.data());
            if (rc == TFS_IO_END) {
                ended = true;
                break;
            }
            if (rc) return rc;
            rc = tfs_io_sources_from_states(nbod, nfil, nseg, nblob, dl, quaternion_based ? 1 : 0, body.data(),
                                            seg.data(), fil_refs.data(), blob_refs.data(), xs.data(),
                                            nblob > 0 ? xb.data() : nullptr, xbod.data());
            if (rc) return rc;
            rc = tfs_flow_set_sources(flow, xs.data(), fseg.data(), nblob > 0 ? xb.data() : nullptr,
                                      nblob > 0 ? fblob.data() : nullptr, xbod.data());
            if (rc) return rc;
            if (advance)  // Euler step with the previous save's velocities (flow_field_evaluator.cu:511-522)
                for (size_t i = 0; i < xf.size(); ++i) xf[i] += dt_save * vf[i];
            if ((rc = tfs_flow_evaluate(flow, nflow, xf.data(), vf.data()))) return rc;
            if ((rc = tfs_io_write_record(loc_path.c_str(), TFS_IO_APPEND | TFS_IO_WITH_STEP, step,
                                          (int64_t)xf.size(), xf.data())))
                return rc;
            if ((rc = tfs_io_write_record(vel_path.c_str(), TFS_IO_APPEND | TFS_IO_WITH_STEP, step,
                                          (int64_t)vf.size(), vf.data())))
                return rc;
            ++done;
            if (saves) *saves = done;
        }
        if (ended && done == 0) break;
        tfs_io_rewind(rd.body);
        tfs_io_rewind(rd.seg);
        tfs_io_rewind(rd.force);
        if (rd.blob) tfs_io_rewind(rd.blob);
    }
    if (saves) *saves = done;
    return TFS_OK;
}

}  // extern "C"
