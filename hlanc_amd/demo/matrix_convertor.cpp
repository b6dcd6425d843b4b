// matrix_convertor.cpp -- mm/pp to bin, as demo/matrix_convertor.cu:12-32.
//   usage: matrix_convertor <mm|pp> <matrix_file> <binary_file>
#include "HLanc.hpp"

int main(int argc, char *argv[])
{
    if (argc != 4 && argc != 5) {
        HLanc::warn("matrix_convertor: convert a sparse matrix into binary csr format.\n");
        HLanc::warn("Usage: %s <mm|pp> <matrix_file> <binary_file>\n", argv[0]);
        return 1;
    }
    HLanc::dcsr_matrix_t m(argv[1], argv[2]);
    HLanc::warn("%10s: %6d\n%10s: %6d\n%10s: %6d\n", "rows", m.get_num_rows(), "columns", m.get_num_columns(), "nnz",
                m.get_nnz());
    m.save_bin(argv[3]);
    HLanc::warn("Success.\n");
    return 0;
}
