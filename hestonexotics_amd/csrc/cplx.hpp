// cplx.hpp -- plain FP64 complex number shared by host and device code.
#pragma once

namespace hexo {

struct Cplx {
    double re, im;
};

}  // namespace hexo
