#pragma once
// output::hdf5_xdmf (output/hdf5_xdmf.hpp) lives with output::hdf5
#include <libmpdata++/output/hdf5.hpp>
