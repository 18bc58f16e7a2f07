// common.h — small utilities with the reference's names (src/common.cpp:5-87).
#pragma once
#include <string>

#include "dataloader.h"
#include "pattern.h"
#include "schedule.h"

double get_wall_time();
void PatternType_printer(PatternType type);
bool is_equal_adj_mat(const int* adj_mat1, const int* adj_mat2, int size);
void GetDataType(DataType& type, const std::string str);
