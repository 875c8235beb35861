// feec/sparse/rainbow.hpp -- Rainbow of FEEC++ (sparse/rainbow.hpp:13-31):
// the greedy multicolouring of a CSR matrix's rows, bit for bit the
// reference's (do_shuffle = false).
#pragma once

#include <string>
#include <vector>

#include "../base/include.hpp"
#include "matcsr.hpp"

struct Rainbow {
    explicit Rainbow(const MatrixCSR&, bool do_shuffle = false);
    void check() const;
    std::string text() const;

    int num_rows;
    int num_colors;
    std::vector<int> F;  // colour of each row
    std::vector<int> B;  // colour offsets (num_colors + 1)
    std::vector<int> R;  // rows grouped by colour
};
