// qtp_kernels.hpp -- gfx950 device code of the qTIP4P/F path-integral step
// (every kernel FP64; layout notes in qtp_common.hpp).
#pragma once
#include "qtp_common.hpp"
#include "qtp_realspace.hpp"
#include "qtp_kspace.hpp"
#include "qtp_pair.hpp"
#include "qtp_step.hpp"
#include "qtp_cells.hpp"
#include "qtp_nn.hpp"
