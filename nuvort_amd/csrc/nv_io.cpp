// The reference's on-disk formats, written and read from host arrays (no
// GPU work): the restart file var.dat (src/output.f90:85-94 data_dump,
// src/initial.f90:151-166 data_restore), the MATLAB stream var%04d.log
// (src/output.f90:47-70 printf), the ts.log line (src/output.f90:23-37) and
// the leaf list of tree_print (src/tree.f90:96-102).
#include "nv_io.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

// Fortran unformatted sequential records: each record's bytes framed by its
// length as a 4-byte integer before and after.
bool put_record(FILE *f, const void *data, size_t bytes) {
    if (bytes > 0x7fffffffu) return false;
    const int32_t len = (int32_t)bytes;
    return std::fwrite(&len, 4, 1, f) == 1 && (bytes == 0 || std::fwrite(data, 1, bytes, f) == bytes) &&
           std::fwrite(&len, 4, 1, f) == 1;
}

// the next record, which must hold `bytes` bytes
bool get_record(FILE *f, void *data, size_t bytes) {
    int32_t a = 0, b = 0;
    if (std::fread(&a, 4, 1, f) != 1 || (size_t)a != bytes) return false;
    if (bytes && std::fread(data, 1, bytes, f) != bytes) return false;
    return std::fread(&b, 4, 1, f) == 1 && a == b;
}

// One qvort as unformatted I/O lays it out (src/cdata.f90:19-29): the
// components in order, unpadded, 248 bytes.
constexpr size_t QVORT_BYTES = 248;

struct Cursor {
    unsigned char *p;
    template <class T>
    void put(T v) {
        std::memcpy(p, &v, sizeof(T));
        p += sizeof(T);
    }
    template <class T>
    T get() {
        T v;
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return v;
    }
};

double at3(const double *a, long r, int k) { return a ? a[3 * r + k] : 0.0; }

}  // namespace

// .true. as the reference's compiler stores a logical(4) in unformatted records
constexpr int32_t FORTRAN_TRUE = 1;

extern "C" int nv_var_write(const char *path, const nv_state *h) {
    if (!path || !h || h->pcount < 0 || h->capacity < h->pcount || (h->pcount > 0 && (!h->x || !h->infront)))
        return NV_ERR_ARG;
    const long n = h->pcount;
    if ((size_t)n * QVORT_BYTES > 0x7fffffffu) return NV_ERR_CAPACITY;   // one record of f
    std::vector<unsigned char> buf((size_t)n * QVORT_BYTES);
    Cursor c{buf.data()};
    for (long r = 0; r < n; ++r) {
        const double *vec[] = {h->x, h->u, h->u1, h->u2, h->u_mf, h->u_sup};
        for (const double *a : vec)
            for (int k = 0; k < 3; ++k) c.put<double>(at3(a, r, k));
        c.put<double>(h->curv ? h->curv[r] : 0.0);
        c.put<double>(h->stretch ? h->stretch[r] : 0.0);
        for (int k = 0; k < 3; ++k) c.put<double>(at3(h->ghosti, r, k));
        for (int k = 0; k < 3; ++k) c.put<double>(at3(h->ghostb, r, k));
        c.put<int32_t>(h->infront[r]);
        c.put<int32_t>(h->behind ? h->behind[r] : 0);
        c.put<int32_t>(h->closest ? h->closest[r] : 0);
        c.put<double>(h->closestd ? h->closestd[r] : 0.0);
        for (int k = 0; k < 3; ++k) c.put<int32_t>(h->wpinned ? h->wpinned[3 * r + k] : 0);
        c.put<int32_t>(h->pinnedi && h->pinnedi[r] ? FORTRAN_TRUE : 0);
        c.put<int32_t>(h->pinnedb && h->pinnedb[r] ? FORTRAN_TRUE : 0);
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) return NV_ERR_ARG;
    const int32_t pc = h->pcount, rc = h->recon_count, it = h->itime;
    const double t = h->t;
    const bool ok = put_record(f, &pc, 4) && put_record(f, &rc, 4) && put_record(f, &it, 4) && put_record(f, &t, 8) &&
                    put_record(f, buf.data(), buf.size());
    return (std::fclose(f) == 0 && ok) ? NV_OK : NV_ERR_ARG;
}

extern "C" int nv_var_read(const char *path, nv_state *h) {
    if (!path || !h) return NV_ERR_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) return NV_ERR_ARG;
    int32_t pc = 0, rc = 0, it = 0;
    double t = 0.0;
    if (!get_record(f, &pc, 4) || !get_record(f, &rc, 4) || !get_record(f, &it, 4) || !get_record(f, &t, 8) ||
        pc < 0) {
        std::fclose(f);
        return NV_ERR_ARG;
    }
    if (h->capacity < pc) {
        std::fclose(f);
        h->pcount = pc;
        return NV_ERR_CAPACITY;
    }
    std::vector<unsigned char> buf((size_t)pc * QVORT_BYTES);
    const bool ok = get_record(f, buf.data(), buf.size());
    std::fclose(f);
    if (!ok) return NV_ERR_ARG;
    Cursor c{buf.data()};
    for (long r = 0; r < pc; ++r) {
        double *vec[] = {h->x, h->u, h->u1, h->u2, h->u_mf, h->u_sup};
        for (double *a : vec)
            for (int k = 0; k < 3; ++k) {
                const double v = c.get<double>();
                if (a) a[3 * r + k] = v;
            }
        const double cv = c.get<double>(), sv = c.get<double>();
        if (h->curv) h->curv[r] = cv;
        if (h->stretch) h->stretch[r] = sv;
        double *gh[] = {h->ghosti, h->ghostb};
        for (double *a : gh)
            for (int k = 0; k < 3; ++k) {
                const double v = c.get<double>();
                if (a) a[3 * r + k] = v;
            }
        const int32_t inf = c.get<int32_t>(), beh = c.get<int32_t>(), clo = c.get<int32_t>();
        const double cd = c.get<double>();
        if (h->infront) h->infront[r] = inf;
        if (h->behind) h->behind[r] = beh;
        if (h->closest) h->closest[r] = clo;
        if (h->closestd) h->closestd[r] = cd;
        for (int k = 0; k < 3; ++k) {
            const int32_t w = c.get<int32_t>();
            if (h->wpinned) h->wpinned[3 * r + k] = w;
        }
        const int32_t pi = c.get<int32_t>(), pb = c.get<int32_t>();
        if (h->pinnedi) h->pinnedi[r] = pi != 0;
        if (h->pinnedb) h->pinnedb[r] = pb != 0;
    }
    h->pcount = pc;
    h->recon_count = rc;
    h->itime = it;
    h->t = t;
    return NV_OK;
}

extern "C" int nv_var_log_write(const char *path, const nv_state *h) {
    if (!path || !h || h->pcount < 0 || (h->pcount > 0 && (!h->x || !h->infront))) return NV_ERR_ARG;
    const long n = h->pcount;
    FILE *f = std::fopen(path, "wb");
    if (!f) return NV_ERR_ARG;
    // stream access: no record markers; t, pcount, then each column
    bool ok = std::fwrite(&h->t, 8, 1, f) == 1 && std::fwrite(&h->pcount, 4, 1, f) == 1;
    std::vector<double> col((size_t)n);
    auto put_vec = [&](const double *a, int k) {
        for (long r = 0; r < n; ++r) col[r] = at3(a, r, k);
        return n == 0 || std::fwrite(col.data(), 8, (size_t)n, f) == (size_t)n;
    };
    for (int k = 0; k < 3 && ok; ++k) ok = put_vec(h->x, k);
    ok = ok && (n == 0 || std::fwrite(h->infront, 4, (size_t)n, f) == (size_t)n);
    for (int k = 0; k < 3 && ok; ++k) ok = put_vec(h->u, k);
    for (int k = 0; k < 3 && ok; ++k) ok = put_vec(h->u_mf, k);
    const double *sc[] = {h->curv, h->stretch};
    for (const double *a : sc) {
        if (!ok) break;
        for (long r = 0; r < n; ++r) col[r] = a ? a[r] : 0.0;
        ok = n == 0 || std::fwrite(col.data(), 8, (size_t)n, f) == (size_t)n;
    }
    return (std::fclose(f) == 0 && ok) ? NV_OK : NV_ERR_ARG;
}

namespace nvio {

// Fortran Iw.m: at least m digits, right-justified in w columns, all
// asterisks when the number does not fit
static void put_i(std::string &s, long v, int w, int m) {
    char b[64];
    std::snprintf(b, sizeof b, "%*.*ld", w, m, v);
    s += (int)std::strlen(b) > w ? std::string(w, '*') : std::string(b);
}

// Fortran Fw.d (the leading zero of |v| < 1 kept when it fits, dropped when
// only it overflows the field, as the processor may)
static void put_f(std::string &s, double v, int w, int d) {
    char b[128];
    std::snprintf(b, sizeof b, "%*.*f", w, d, v);
    std::string t(b);
    if ((int)t.size() > w && std::fabs(v) < 1.0) {
        const size_t z = t.find("0.");
        if (z != std::string::npos) t.erase(z, 1);
    }
    s += (int)t.size() > w ? std::string(w, '*') : t;
}

std::string ts_header() {
    return " %--var--------t-------pcount--------recon-----wall_recon----avg_d-----"
           "                   length--------maxu---------maxdu------curv------removed-----tree_eval\n";
}

std::string ts_line(int shots, const nv_diag &d, double delta) {
    std::string s;
    put_i(s, shots > 0 ? d.itime / shots : 0, 6, 4);
    put_f(s, d.t, 13, 7);
    put_i(s, d.active, 10, 1);
    put_i(s, d.recon_count, 13, 1);
    put_i(s, d.wall_recon_count, 13, 1);
    put_f(s, d.avg_sep / delta, 7, 4);
    put_f(s, d.total_length, 13, 6);
    put_f(s, d.maxu, 13, 5);
    put_f(s, d.maxdu, 13, 5);
    put_f(s, d.kappa_bar, 10, 2);
    put_i(s, d.remove_count, 13, 1);
    put_i(s, (long)d.tree_eval, 13, 1);
    return s + "\n";
}

std::string tree_leaf_line(const nv_tree_node &nd) {
    char b[256];
    std::snprintf(b, sizeof b, " %.17E %.17E %.17E %.17E %.17E %.17E %d\n", nd.pos[0], nd.pos[0] + nd.width, nd.pos[1],
                  nd.pos[1] + nd.width, nd.pos[2], nd.pos[2] + nd.width, nd.pcount);
    return b;
}

}  // namespace nvio
