/* orc_math.h -- TEST INFRASTRUCTURE ONLY: the oracle's FP64 log/exp. */
#ifndef ORC_MATH_H
#define ORC_MATH_H
double orc_log_cr(double x);
double orc_exp_cr(double x);
double orc_log_glibc(double x);
double orc_exp_glibc(double x);
unsigned long long orc_math_unresolved(void);
#endif
