// feature_if.h -- interface of a runnable feature, as the reference's
// feature_if (include/feature_if.h:27-33).
#ifndef EMP_FEATURE_IF_H
#define EMP_FEATURE_IF_H

namespace edge_matching_puzzle {

class feature_if {
  public:
    virtual void run() = 0;
    virtual ~feature_if() = default;
};

}  // namespace edge_matching_puzzle

#endif
