/*
 * reduce.h -- reduction operators accepted by PSReduce.
 * The enumerator order is ABI: PS_MAX=0, PS_MIN=1, PS_SUM=2, PS_PROD=3
 * (reference include/physis/reduce.h:16-21).
 */
#ifndef PHYSIS_REDUCE_H_
#define PHYSIS_REDUCE_H_

#ifdef __cplusplus
extern "C" {
#endif

enum PSReduceOp {
  PS_MAX,
  PS_MIN,
  PS_SUM,
  PS_PROD
};

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_REDUCE_H_ */
