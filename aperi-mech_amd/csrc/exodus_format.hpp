// Exodus II results file in the netCDF classic format, 64-bit-offset variant (CDF-2),
// written directly (no netCDF or Exodus library in this build).  Host-only C++.
//
// netCDF classic layout (all integers big-endian):
//   "CDF\x02"  numrecs(u32)  dim_list  gatt_list  var_list  | non-record data | records
//   dim_list  = 0x0A n {name, length}            (length 0 = the record dimension)
//   gatt_list = 0x0C n {name, type, n, values}   (values padded to 4 bytes)
//   var_list  = 0x0B n {name, ndims, dimids, vatt_list, type, vsize(u32), begin(u64)}
//   name = length(u32) + bytes padded to 4; an absent list is two zero words.
// A record holds one slice of every record variable in declaration order.
//
// Exodus II conventions used (the names the Exodus library and exodiff read): one
// element block; coordx/coordy/coordz; connect1 (1-based) with its elem_type attribute;
// time_whole; vals_nod_var<i>(time_step, num_nodes); vals_elem_var<i>eb1(time_step,
// num_el_in_blk1); name_nod_var / name_elem_var; elem_var_tab; the global attributes
// api_version, version, floating_point_word_size (8), file_size (1),
// maximum_name_length (32), int64_status (0) and title.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace aphip {
namespace exo {

enum NcType : uint32_t { NC_CHAR = 2, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };

class Bytes {
   public:
    void u32(uint32_t v) {
        for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void u64(uint64_t v) {
        for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void f64(double d) {
        uint64_t v;
        std::memcpy(&v, &d, 8);
        u64(v);
    }
    void f32(float f) {
        uint32_t v;
        std::memcpy(&v, &f, 4);
        u32(v);
    }
    void text(const std::string &s, size_t width) {  // exactly width bytes, zero filled
        for (size_t i = 0; i < width; ++i) b.push_back(i < s.size() ? (uint8_t)s[i] : 0);
    }
    void pad4() {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string &s) {
        u32((uint32_t)s.size());
        text(s, s.size());
        pad4();
    }
    std::vector<uint8_t> b;
};

struct Mesh {
    std::string title;
    int64_t n_nodes = 0;
    const double *x = nullptr, *y = nullptr, *z = nullptr;  // n_nodes each
    int64_t n_elem = 0;
    int npe = 0;
    const int32_t *conn = nullptr;  // n_elem * npe, 0-based
    std::string elem_type;          // "TETRA4", "HEX8"
    std::string block_name = "block_1";
};

constexpr int kLenName = 33;  // maximum_name_length 32 + terminator

class Writer {
   public:
    Writer(const std::string &path, const Mesh &m, const std::vector<std::string> &node_vars,
           const std::vector<std::string> &elem_vars)
        : m_nn(m.n_nodes), m_ne(m.n_elem), m_nnv((int)node_vars.size()), m_nev((int)elem_vars.size()) {
        if (m.n_nodes <= 0 || m.n_elem < 0 || (m.n_elem > 0 && (m.npe <= 0 || !m.conn)) || !m.x || !m.y || !m.z)
            throw std::runtime_error("results file: invalid mesh");
        for (const auto &v : node_vars)
            if (v.size() >= (size_t)kLenName) throw std::runtime_error("results file: variable name too long");
        for (const auto &v : elem_vars)
            if (v.size() >= (size_t)kLenName) throw std::runtime_error("results file: variable name too long");
        m_fp = std::fopen(path.c_str(), "wb");
        if (!m_fp) throw std::runtime_error("results file: cannot create " + path);

        // dimensions (ids in this order)
        struct Dim {
            const char *name;
            int64_t len;
        };
        const int64_t nev = m_nev, nnv = m_nnv;
        std::vector<Dim> dims = {{"len_string", 33}, {"len_line", 81}, {"four", 4}, {"len_name", kLenName},
                                 {"time_step", 0}, {"num_dim", 3}, {"num_nodes", m.n_nodes},
                                 {"num_elem", m.n_elem}, {"num_el_blk", 1}, {"num_el_in_blk1", m.n_elem},
                                 {"num_nod_per_el1", m.npe}};
        if (nnv) dims.push_back({"num_nod_var", nnv});
        if (nev) dims.push_back({"num_elem_var", nev});
        auto did = [&](const char *n) {
            for (size_t i = 0; i < dims.size(); ++i)
                if (std::strcmp(dims[i].name, n) == 0) return (uint32_t)i;
            throw std::logic_error("results file: unknown dimension");
        };

        // variables: non-record data first, then the record variables
        struct Attr {
            std::string name, text;
        };
        struct Var {
            std::string name;
            std::vector<uint32_t> dims;
            NcType type;
            Attr attr;
            Bytes data;  // non-record variables: their bytes (padded)
            uint64_t vsize = 0, begin = 0;
            bool record = false;
        };
        std::vector<Var> vars;
        auto add = [&](const std::string &n, std::vector<uint32_t> d, NcType t, bool rec) -> Var & {
            vars.push_back(Var{n, std::move(d), t, Attr{}, Bytes{}, 0, 0, rec});
            return vars.back();
        };
        {
            Var &v = add("eb_status", {did("num_el_blk")}, NC_INT, false);
            v.data.u32(1);
        }
        {
            Var &v = add("eb_prop1", {did("num_el_blk")}, NC_INT, false);
            v.attr = Attr{"name", "ID"};
            v.data.u32(1);
        }
        {
            Var &v = add("eb_names", {did("num_el_blk"), did("len_name")}, NC_CHAR, false);
            v.data.text(m.block_name, kLenName);
            v.data.pad4();
        }
        const double *xyz[3] = {m.x, m.y, m.z};
        const char *cn[3] = {"coordx", "coordy", "coordz"};
        for (int d = 0; d < 3; ++d) {
            Var &v = add(cn[d], {did("num_nodes")}, NC_DOUBLE, false);
            for (int64_t i = 0; i < m.n_nodes; ++i) v.data.f64(xyz[d][i]);
        }
        {
            Var &v = add("coor_names", {did("num_dim"), did("len_name")}, NC_CHAR, false);
            for (const char *c : {"x", "y", "z"}) v.data.text(c, kLenName);
            v.data.pad4();
        }
        {
            Var &v = add("connect1", {did("num_el_in_blk1"), did("num_nod_per_el1")}, NC_INT, false);
            v.attr = Attr{"elem_type", m.elem_type};
            for (int64_t i = 0; i < m.n_elem * m.npe; ++i) {
                if (m.conn[i] < 0 || m.conn[i] >= m.n_nodes) throw std::runtime_error("results file: bad connectivity");
                v.data.u32((uint32_t)(m.conn[i] + 1));
            }
        }
        if (nnv) {
            Var &v = add("name_nod_var", {did("num_nod_var"), did("len_name")}, NC_CHAR, false);
            for (const auto &n : node_vars) v.data.text(n, kLenName);
            v.data.pad4();
        }
        if (nev) {
            Var &v = add("name_elem_var", {did("num_elem_var"), did("len_name")}, NC_CHAR, false);
            for (const auto &n : elem_vars) v.data.text(n, kLenName);
            v.data.pad4();
            Var &t = add("elem_var_tab", {did("num_el_blk"), did("num_elem_var")}, NC_INT, false);
            for (int64_t i = 0; i < nev; ++i) t.data.u32(1);
        }
        add("time_whole", {did("time_step")}, NC_DOUBLE, true).vsize = 8;
        for (int64_t i = 0; i < nnv; ++i)
            add("vals_nod_var" + std::to_string(i + 1), {did("time_step"), did("num_nodes")}, NC_DOUBLE, true).vsize =
                8 * (uint64_t)m.n_nodes;
        for (int64_t i = 0; i < nev; ++i)
            add("vals_elem_var" + std::to_string(i + 1) + "eb1", {did("time_step"), did("num_el_in_blk1")}, NC_DOUBLE,
                true)
                .vsize = 8 * (uint64_t)m.n_elem;
        for (auto &v : vars)
            if (!v.record) v.vsize = v.data.b.size();

        // header: sized once with zero offsets, then written with the real ones
        auto header = [&]() {
            Bytes h;
            h.text("CDF", 3);
            h.b.push_back(2);
            h.u32(0);  // numrecs, updated after every record
            h.u32(0x0A);
            h.u32((uint32_t)dims.size());
            for (const auto &d : dims) {
                h.name(d.name);
                h.u32((uint32_t)d.len);
            }
            struct GAttr {
                const char *name;
                NcType type;
                double value;
                std::string text;
            };
            const std::vector<GAttr> ga = {{"api_version", NC_FLOAT, 8.03, ""},
                                           {"version", NC_FLOAT, 8.03, ""},
                                           {"floating_point_word_size", NC_INT, 8, ""},
                                           {"file_size", NC_INT, 1, ""},
                                           {"maximum_name_length", NC_INT, 32, ""},
                                           {"int64_status", NC_INT, 0, ""},
                                           {"title", NC_CHAR, 0, m.title}};
            h.u32(0x0C);
            h.u32((uint32_t)ga.size());
            for (const auto &a : ga) {
                h.name(a.name);
                h.u32(a.type);
                if (a.type == NC_CHAR) {
                    h.u32((uint32_t)a.text.size());
                    h.text(a.text, a.text.size());
                    h.pad4();
                } else if (a.type == NC_FLOAT) {
                    h.u32(1);
                    h.f32((float)a.value);
                } else {
                    h.u32(1);
                    h.u32((uint32_t)(int32_t)a.value);
                }
            }
            h.u32(0x0B);
            h.u32((uint32_t)vars.size());
            for (const auto &v : vars) {
                h.name(v.name);
                h.u32((uint32_t)v.dims.size());
                for (uint32_t d : v.dims) h.u32(d);
                if (v.attr.name.empty()) {
                    h.u32(0);
                    h.u32(0);
                } else {
                    h.u32(0x0C);
                    h.u32(1);
                    h.name(v.attr.name);
                    h.u32(NC_CHAR);
                    h.u32((uint32_t)v.attr.text.size());
                    h.text(v.attr.text, v.attr.text.size());
                    h.pad4();
                }
                h.u32(v.type);
                h.u32((uint32_t)v.vsize);
                h.u64(v.begin);
            }
            return h;
        };
        uint64_t off = header().b.size();
        for (auto &v : vars)
            if (!v.record) {
                v.begin = off;
                off += v.vsize;
            }
        m_record_begin = off;
        for (auto &v : vars)
            if (v.record) {
                v.begin = off;
                off += v.vsize;
            }
        m_record_size = off - m_record_begin;
        Bytes h = header();
        for (const auto &v : vars)
            if (!v.record) h.b.insert(h.b.end(), v.data.b.begin(), v.data.b.end());
        if (h.b.size() != m_record_begin) throw std::logic_error("results file: header layout");
        put(h.b);
    }

    Writer(const Writer &) = delete;
    Writer &operator=(const Writer &) = delete;
    ~Writer() { close(); }

    // One time step: node_vals[i] holds n_nodes values of node variable i, elem_vals[j]
    // n_elem values of element variable j (declaration order).
    void write_step(double time, const std::vector<const double *> &node_vals,
                    const std::vector<const double *> &elem_vals) {
        if (!m_fp) throw std::runtime_error("results file is closed");
        if ((int)node_vals.size() != m_nnv || (int)elem_vals.size() != m_nev)
            throw std::runtime_error("results file: wrong number of variables");
        Bytes r;
        r.b.reserve((size_t)m_record_size);
        r.f64(time);
        for (const double *p : node_vals)
            for (int64_t i = 0; i < m_nn; ++i) r.f64(p[i]);
        for (const double *p : elem_vals)
            for (int64_t i = 0; i < m_ne; ++i) r.f64(p[i]);
        if (std::fseek(m_fp, (long)(m_record_begin + m_records * m_record_size), SEEK_SET) != 0)
            throw std::runtime_error("results file: seek failed");
        put(r.b);
        ++m_records;
        Bytes n;
        n.u32((uint32_t)m_records);
        if (std::fseek(m_fp, 4, SEEK_SET) != 0) throw std::runtime_error("results file: seek failed");
        put(n.b);
        std::fflush(m_fp);
    }

    int64_t records() const { return m_records; }

    void close() {
        if (m_fp) std::fclose(m_fp);
        m_fp = nullptr;
    }

   private:
    void put(const std::vector<uint8_t> &b) {
        if (std::fwrite(b.data(), 1, b.size(), m_fp) != b.size()) throw std::runtime_error("results file: write failed");
    }
    FILE *m_fp = nullptr;
    int64_t m_nn, m_ne;
    int m_nnv, m_nev;
    uint64_t m_record_begin = 0, m_record_size = 0;
    int64_t m_records = 0;
};

}  // namespace exo
}  // namespace aphip
