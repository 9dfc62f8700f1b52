// th_io.cpp -- the reference's field files, host side only (plain host arrays in the reference
// layout; no GPU is touched, so checkpoints can be written and read anywhere).
//
// Format (field_out.d written by Output_field, field_in.d read with field_input_proc = 1):
// ASCII; each snapshot is one "%% Time = t" line followed, per component, by the
// N1 x N2 x (N3/2+1) complex values in (kx, ky, kz) row-major order, each printed as
// "(re,im) " with 12 significant digits, and a blank line after each component.
//   writer: IncFluid::Output_field output_field.cc:43-58 -> CVF::CV_output Cvf.cc:715-720 ->
//           Write_data_MPI array_basic.cc:814-858 (ostream << complex<DP> << " ", then endl endl),
//           stream precision MY_PRECISION = 12 (extern_def_vars.h:78, file_operations.cc:230)
//   reader: IncFluid::Init_cond init_cond_field.cc:47-86 -> CVF::CV_input Cvf.cc:739-744 ->
//           Read_data_MPI array_basic.cc:294-335 (istream >> complex<DP>, so "(re,im)", "(re)" and
//           a bare "re" are all values)
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <thread>
#include <vector>
#include "../../include/tarang_hip.h"
#include "th_internal.hpp"

namespace {

// One value as std::ostream with precision 12 prints a complex<double>: "(re,im)" with each part
// in the %g style at 12 significant digits.
int put_value(char* out, double re, double im)
{
    return snprintf(out, 64, "(%.12g,%.12g) ", re, im);
}

template <typename R> void format_range(const R* a, size_t i0, size_t i1, std::string& out)
{
    char buf[64];
    out.clear();
    out.reserve((i1 - i0) * 40);
    for (size_t i = i0; i < i1; i++) {
        int n = put_value(buf, (double)a[2 * i], (double)a[2 * i + 1]);
        out.append(buf, (size_t)n);
    }
}

// Formats n values with up to nthr host threads, in order.
template <typename R> bool write_component(FILE* f, const R* a, size_t n, int nthr)
{
    const size_t chunk = 1 << 20;
    std::vector<std::string> parts((size_t)nthr);
    for (size_t base = 0; base < n; base += chunk * nthr) {
        std::vector<std::thread> th;
        for (int k = 0; k < nthr; k++) {
            size_t i0 = base + (size_t)k * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
            if (i0 >= n) {
                parts[k].clear();
                continue;
            }
            th.emplace_back([&, k, i0, i1] { format_range(a, i0, i1, parts[k]); });
        }
        for (auto& t : th) t.join();
        for (int k = 0; k < nthr; k++)
            if (!parts[k].empty() && fwrite(parts[k].data(), 1, parts[k].size(), f) != parts[k].size()) return false;
    }
    return fputs("\n\n", f) >= 0;
}

struct Reader {
    const char* p;
    const char* end;
    void skip_space()
    {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
    }
    // "%% ..." line: returns true (and the time when the line is "%% Time = t")
    bool header(double* time)
    {
        if (end - p < 2 || p[0] != '%' || p[1] != '%') return false;
        const char* eol = (const char*)memchr(p, '\n', (size_t)(end - p));
        if (!eol) eol = end;
        std::string line(p, eol);
        size_t k = line.find("Time");
        if (k != std::string::npos) {
            size_t e = line.find('=', k);
            if (e != std::string::npos && time) *time = strtod(line.c_str() + e + 1, nullptr);
        }
        p = eol;
        return true;
    }
    bool number(double* v)
    {
        char* q;
        *v = strtod(p, &q);
        if (q == p) return false;
        p = q;
        return true;
    }
    // one complex value in one of the forms istream >> complex accepts
    bool value(double* re, double* im)
    {
        skip_space();
        if (p >= end) return false;
        *im = 0.0;
        if (*p != '(') return number(re);
        p++;
        skip_space();
        if (!number(re)) return false;
        skip_space();
        if (p < end && *p == ',') {
            p++;
            skip_space();
            if (!number(im)) return false;
            skip_space();
        }
        if (p >= end || *p != ')') return false;
        p++;
        return true;
    }
};

bool load_file(const char* path, std::vector<char>& buf)
{
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    char tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    fclose(f);
    buf.push_back('\0');
    return true;
}

bool args_ok(const char* path, const int N[3], int precision, int ncomp, const void* fields)
{
    return path && N && N[0] > 0 && N[1] > 0 && N[2] > 1 && ncomp > 0 && fields &&
           (precision == TH_FP64 || precision == TH_FP32_MIXED);
}

}  // namespace

extern "C" int th_field_write(const char* path, const int N[3], int precision, int ncomp, const void* const fields[],
                              double time, int append)
{
    if (!args_ok(path, N, precision, ncomp, fields)) {
        th::set_error("th_field_write: bad args");
        return TH_E_ARG;
    }
    for (int c = 0; c < ncomp; c++)
        if (!fields[c]) {
            th::set_error("th_field_write: component %d is null", c);
            return TH_E_ARG;
        }
    FILE* f = fopen(path, append ? "ab" : "wb");
    if (!f) {
        th::set_error("th_field_write: cannot open %s", path);
        return TH_E_IO;
    }
    const size_t n = (size_t)N[0] * N[1] * (N[2] / 2 + 1);
    unsigned hw = std::thread::hardware_concurrency();
    const int nthr = hw == 0 ? 1 : (hw > 16 ? 16 : (int)hw);
    char head[96];
    snprintf(head, sizeof head, "%%%% Time = %.12g\n", time);
    bool ok = fputs(head, f) >= 0;
    for (int c = 0; c < ncomp && ok; c++)
        ok = precision == TH_FP64 ? write_component(f, (const double*)fields[c], n, nthr)
                                  : write_component(f, (const float*)fields[c], n, nthr);
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        th::set_error("th_field_write: write to %s failed", path);
        return TH_E_IO;
    }
    return TH_OK;
}

extern "C" int th_field_read(const char* path, const int N[3], int precision, int ncomp, void* const fields[],
                             int snapshot, double* time)
{
    if (!args_ok(path, N, precision, ncomp, fields) || snapshot < 0) {
        th::set_error("th_field_read: bad args");
        return TH_E_ARG;
    }
    for (int c = 0; c < ncomp; c++)
        if (!fields[c]) {
            th::set_error("th_field_read: component %d is null", c);
            return TH_E_ARG;
        }
    std::vector<char> buf;
    if (!load_file(path, buf)) {
        th::set_error("th_field_read: cannot open %s", path);
        return TH_E_IO;
    }
    Reader rd{buf.data(), buf.data() + buf.size() - 1};
    const size_t n = (size_t)N[0] * N[1] * (N[2] / 2 + 1);
    double t = NAN, re, im;
    for (int s = 0; s <= snapshot; s++) {
        rd.skip_space();
        while (rd.header(&t)) rd.skip_space();
        for (int c = 0; c < ncomp; c++)
            for (size_t i = 0; i < n; i++) {
                if (!rd.value(&re, &im)) {
                    th::set_error("th_field_read: %s: snapshot %d, component %d: expected %zu values, value %zu is "
                                  "missing or malformed",
                                  path, s, c, n, i);
                    return TH_E_IO;
                }
                if (s < snapshot) continue;
                if (precision == TH_FP64) {
                    ((double*)fields[c])[2 * i] = re;
                    ((double*)fields[c])[2 * i + 1] = im;
                } else {
                    ((float*)fields[c])[2 * i] = (float)re;
                    ((float*)fields[c])[2 * i + 1] = (float)im;
                }
            }
    }
    if (time) *time = t;
    return TH_OK;
}
