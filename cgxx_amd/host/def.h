// Scalar type of the CG path.  Same as the reference (def.h:6): the
// solver's public interface is FP64 throughout.
#ifndef CGHIP_HOST_DEF_H
#define CGHIP_HOST_DEF_H

using floatType = double;

#endif
