// motif_generator.h — all connected patterns of a given size, one per
// isomorphism class, in the reference's order (include/motif_generator.h:1-13).
#pragma once
#include <vector>

#include "pattern.h"

class MotifGenerator {
public:
    MotifGenerator(int input_size) : size(input_size) {}
    std::vector<Pattern> generate();

private:
    int size;
};
