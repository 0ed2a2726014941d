/*
 * types.h -- primitive point-type tags (reference include/physis/types.h:15-22).
 */
#ifndef PHYSIS_TYPES_H_
#define PHYSIS_TYPES_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef int PSType;
enum PSPrimitiveType {
  PS_INT = 0,
  PS_LONG = 1,
  PS_FLOAT = 2,
  PS_DOUBLE = 3,
  PS_USER = 4
};

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_TYPES_H_ */
