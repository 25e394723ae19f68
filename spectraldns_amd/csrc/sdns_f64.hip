// Double-precision instantiation of every pass kernel and launcher.
#include "sdns_launch_impl.hpp"
SDNS_INSTANTIATE(double)
