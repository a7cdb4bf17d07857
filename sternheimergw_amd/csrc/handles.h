// Opaque handle types of the C ABI (include/sgw.h), shared by the C-ABI sources.
#pragma once

#include <memory>

#include "bicgstab.h"
#include "sgw.h"
#include "sgw_internal.h"

struct sgw_operator {
  std::unique_ptr<sgw::Operator> op;
};

struct sgw_batch_solver {
  sgw_operator *op;
  sgw_batch_config cfg;
  std::unique_ptr<sgw::BatchBicgstab> engine;
};
