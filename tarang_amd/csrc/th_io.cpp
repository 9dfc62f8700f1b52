// th_io.cpp -- the reference's field files, host side only (plain host arrays in the reference
// layout; no GPU is touched, so checkpoints can be written and read anywhere).
//
// Format (field_out.d written by Output_field, field_in.d read with field_input_proc = 1):
// ASCII; each snapshot is one "%% Time = t" line followed, per component, by the
// N1 x N2 x (N3/2+1) complex values in (kx, ky, kz) row-major order, each printed as
// "(re,im) " in fixed notation with 12 decimals, and a blank line after each component.
//   writer: IncFluid::Output_field output_field.cc:43-58 -> CVF::CV_output Cvf.cc:715-720 ->
//           Write_data_MPI array_basic.cc:814-858 (ostream << complex<DP> << " ", then endl endl),
//           on a stream set to ios::fixed with precision MY_PRECISION = 12 (extern_def_vars.h:78,
//           file_operations.cc:229-230); the time line on the same stream
// Reduced files (field_out_reduced.d, field_input_proc = 2) hold the modes of a smaller grid; the reference
// leaves that stream at its defaults (6 significant digits) and prints the values with no separator
// (th_field_write_reduced / th_field_read_reduced below).
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

// One value as std::ostream in fixed notation with precision 12 prints a complex<double>: "(re,im)", each
// part with 12 decimals (the complex inserter formats both parts with the stream's flags and precision).
int put_value(char* out, size_t cap, double re, double im)
{
    return snprintf(out, cap, "(%.12f,%.12f) ", re, im);
}

template <typename R> void format_range(const R* a, size_t i0, size_t i1, std::string& out)
{
    char buf[64];
    out.clear();
    out.reserve((i1 - i0) * 40);
    for (size_t i = i0; i < i1; i++) {
        const double re = (double)a[2 * i], im = (double)a[2 * i + 1];
        int n = put_value(buf, sizeof buf, re, im);
        if (n >= (int)sizeof buf) {   // |value| >= 1e30 or so: the fixed form has many integer digits
            std::vector<char> big((size_t)n + 1);
            put_value(big.data(), big.size(), re, im);
            out.append(big.data(), (size_t)n);
        } else {
            out.append(buf, (size_t)n);
        }
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

// Reduced-size files (field_out_reduced.d, field_input_proc = 2).  The file holds the Nr1 x Nr2 x (Nr3/2+1)
// modes of the reduced grid in its own layout: x rows kx = 0..Nr1/2, then -Nr1/2+1..-1 (FOUR; SCFT: kx =
// 0..Nr1-1), each row ky = 0..Nr2/2 then -Nr2/2+1..-1, each with kz = 0..Nr3/2.  row_of / col_of give the row
// and column of the full N grid that reduced row i / column j come from.
int red_row(int basis, int i, int Nr1, int N1) { return (basis == TH_FOUR && i > Nr1 / 2) ? N1 - Nr1 + i : i; }
int red_col(int j, int Nr2, int N2) { return j > Nr2 / 2 ? N2 - Nr2 + j : j; }

bool reduced_ok(const int N[3], const int Nr[3], int basis)
{
    if (!Nr || (basis != TH_FOUR && basis != TH_SCFT)) return false;
    for (int d = 0; d < 3; d++)
        if (Nr[d] < 2 || Nr[d] > N[d] || Nr[d] % 2) return false;
    return true;
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
    std::vector<char> head(400);
    snprintf(head.data(), head.size(), "%%%% Time = %.12f\n", time);
    bool ok = fputs(head.data(), f) >= 0;
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

// Output_field_reduced output_field.cc:505-520 -> CV_output(file, Nreduced) Cvf.cc:732-737 -> Write_data_MPI(...,
// Nreduced, ...) array_basic.cc:997-1170 on one process: the modes of the reduced grid, each printed as
// ostream << complex<DP> with no separator (unlike the full writer's " "), then endl endl, on a stream left at
// its default format (field_out_reduced_file gets no precision or floatfield, file_operations.cc:174-236), so
// the time and the values carry 6 significant digits.
extern "C" int th_field_write_reduced(const char* path, const int N[3], const int Nr[3], int basis, int precision,
                                      int ncomp, const void* const fields[], double time, int append)
{
    if (!args_ok(path, N, precision, ncomp, fields) || !reduced_ok(N, Nr, basis)) {
        th::set_error("th_field_write_reduced: bad args (2 <= Nr[d] <= N[d], Nr[d] even, basis FOUR or SCFT)");
        return TH_E_ARG;
    }
    for (int c = 0; c < ncomp; c++)
        if (!fields[c]) {
            th::set_error("th_field_write_reduced: component %d is null", c);
            return TH_E_ARG;
        }
    FILE* f = fopen(path, append ? "ab" : "wb");
    if (!f) {
        th::set_error("th_field_write_reduced: cannot open %s", path);
        return TH_E_IO;
    }
    const long H = N[2] / 2 + 1, Hr = Nr[2] / 2 + 1;
    char head[96], buf[64];
    snprintf(head, sizeof head, "%%%% Time = %g\n", time);
    bool ok = fputs(head, f) >= 0;
    std::string line;
    for (int c = 0; c < ncomp && ok; c++) {
        for (int i = 0; i < Nr[0] && ok; i++) {
            const long row = red_row(basis, i, Nr[0], N[0]);
            line.clear();
            for (int j = 0; j < Nr[1]; j++) {
                const long base = (row * N[1] + red_col(j, Nr[1], N[1])) * H;
                for (long k = 0; k < Hr; k++) {
                    const long e = base + k;
                    const double re = precision == TH_FP64 ? ((const double*)fields[c])[2 * e] : ((const float*)fields[c])[2 * e];
                    const double im = precision == TH_FP64 ? ((const double*)fields[c])[2 * e + 1]
                                                           : ((const float*)fields[c])[2 * e + 1];
                    int n = snprintf(buf, sizeof buf, "(%g,%g)", re, im);
                    line.append(buf, (size_t)n);
                }
            }
            ok = fwrite(line.data(), 1, line.size(), f) == line.size();
        }
        ok = ok && fputs("\n\n", f) >= 0;
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        th::set_error("th_field_write_reduced: write to %s failed", path);
        return TH_E_IO;
    }
    return TH_OK;
}

// Init_cond_reduced init_cond_field.cc:292-330 -> CV_input(file, Nreduced) Cvf.cc:747-752 -> Read_data_MPI(...,
// Nreduced, ...) array_basic.cc:388-590 on one process: the reduced grid's modes into the full arrays (every
// other mode 0).  The Nyquist planes of the reduced grid land on both sides, as there: the
// ky = Nr2/2 column also at ky = -Nr2/2 (row N2 - Nr2/2), and (FOUR) the kx = Nr1/2 row also at kx = -Nr1/2
// (row N1 - Nr1/2).
extern "C" int th_field_read_reduced(const char* path, const int N[3], const int Nr[3], int basis, int precision,
                                     int ncomp, void* const fields[], int snapshot, double* time)
{
    if (!args_ok(path, N, precision, ncomp, fields) || !reduced_ok(N, Nr, basis) || snapshot < 0) {
        th::set_error("th_field_read_reduced: bad args (2 <= Nr[d] <= N[d], Nr[d] even, basis FOUR or SCFT)");
        return TH_E_ARG;
    }
    for (int c = 0; c < ncomp; c++)
        if (!fields[c]) {
            th::set_error("th_field_read_reduced: component %d is null", c);
            return TH_E_ARG;
        }
    std::vector<char> buf;
    if (!load_file(path, buf)) {
        th::set_error("th_field_read_reduced: cannot open %s", path);
        return TH_E_IO;
    }
    Reader rd{buf.data(), buf.data() + buf.size() - 1};
    const long H = N[2] / 2 + 1, Hr = Nr[2] / 2 + 1;
    const size_t n = (size_t)N[0] * N[1] * H, es = precision == TH_FP64 ? 16 : 8;
    auto put = [&](void* a, long e, double re, double im) {
        if (precision == TH_FP64) {
            ((double*)a)[2 * e] = re;
            ((double*)a)[2 * e + 1] = im;
        } else {
            ((float*)a)[2 * e] = (float)re;
            ((float*)a)[2 * e + 1] = (float)im;
        }
    };
    double t = NAN, re, im;
    for (int s = 0; s <= snapshot; s++) {
        rd.skip_space();
        while (rd.header(&t)) rd.skip_space();
        for (int c = 0; c < ncomp; c++) {
            if (s == snapshot) memset(fields[c], 0, n * es);
            for (int i = 0; i < Nr[0]; i++)
                for (int j = 0; j < Nr[1]; j++)
                    for (long k = 0; k < Hr; k++) {
                        if (!rd.value(&re, &im)) {
                            th::set_error("th_field_read_reduced: %s: snapshot %d, component %d: reduced mode (%d, %d, %ld) "
                                          "is missing or malformed",
                                          path, s, c, i, j, k);
                            return TH_E_IO;
                        }
                        if (s < snapshot) continue;
                        const int rows[2] = {red_row(basis, i, Nr[0], N[0]),
                                             (basis == TH_FOUR && i == Nr[0] / 2) ? N[0] - Nr[0] / 2 : -1};
                        const int cols[2] = {red_col(j, Nr[1], N[1]), j == Nr[1] / 2 ? N[1] - Nr[1] / 2 : -1};
                        for (int a = 0; a < 2; a++)
                            for (int b = 0; b < 2; b++)
                                if (rows[a] >= 0 && cols[b] >= 0) put(fields[c], ((long)rows[a] * N[1] + cols[b]) * H + k, re, im);
                    }
        }
    }
    if (time) *time = t;
    return TH_OK;
}
