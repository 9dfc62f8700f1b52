// Field-file writer and reader built only from the C++ standard library operations the reference
// uses for field_out.d / field_in.d: a stream set to ios::fixed with precision 12 (MY_PRECISION,
// file_operations.cc:229-230), "%% Time = " << t << endl, then per component every value as
// stream << std::complex<double> << " "  and endl endl (Write_data_MPI array_basic.cc:814-858); reading is
// stream >> std::complex<double> (Read_data_MPI array_basic.cc:294-335).  The reduced file
// (field_out_reduced.d) is a stream left at its defaults, every value as  stream << std::complex<double>
// with no separator, and endl endl per component (Write_data_MPI(..., Nreduced, ...) array_basic.cc:997-1170);
// the values arrive here already in file order.  Values travel to and from this program as raw doubles.
//   ref_field_io write    <values.bin> <out.txt> <n_values_per_comp> <ncomp> <time>
//   ref_field_io writered <values.bin> <out.txt> <n_values_per_comp> <ncomp> <time>
//   ref_field_io read     <in.txt> <values.bin> <n_values_per_comp> <ncomp>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

int main(int argc, char** argv)
{
    if (argc < 6) return 2;
    const std::string mode = argv[1];
    const long n = atol(argv[4]);
    const int ncomp = atoi(argv[5]);
    std::vector<double> v(2 * n * ncomp);
    if (mode == "write" || mode == "writered") {
        const bool red = mode == "writered";
        FILE* f = fopen(argv[2], "rb");
        if (!f || fread(v.data(), sizeof(double), v.size(), f) != v.size()) return 3;
        fclose(f);
        std::ofstream out(argv[3]);
        if (!red) {
            out.setf(std::ios::fixed, std::ios::floatfield);
            out.precision(12);
        }
        out << "%% Time = " << atof(argv[6]) << std::endl;
        for (int c = 0; c < ncomp; c++) {
            for (long i = 0; i < n; i++) {
                out << std::complex<double>(v[2 * (c * n + i)], v[2 * (c * n + i) + 1]);
                if (!red) out << " ";
            }
            out << std::endl << std::endl;
        }
        return out.good() ? 0 : 4;
    }
    std::ifstream in(argv[2]);
    std::string line;
    std::getline(in, line);                 // the "%% Time" line
    for (long k = 0; k < n * ncomp; k++) {
        std::complex<double> z;
        in >> z;
        if (!in) return 5;
        v[2 * k] = z.real();
        v[2 * k + 1] = z.imag();
    }
    FILE* f = fopen(argv[3], "wb");
    if (!f || fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) return 6;
    fclose(f);
    return 0;
}
