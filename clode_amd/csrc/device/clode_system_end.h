// Ends one system of a batched program.
//
// A batched program compiles the steppers, observers and drivers once per
// system, each copy inside its own namespace and with its own dimensions,
// stepper and observer.  This header removes every macro those per-system
// headers (and the system's defines) introduce, so the next system's copy is
// compiled from a clean slate.
#undef CLODE_STEPPER
#undef CLODE_OBSERVER
#undef CLODE_EVAR
#undef CLODE_FVAR
#undef N_VAR
#undef N_PAR
#undef N_AUX
#undef N_WIENER
#undef N_STORE_EVENTS
// clode_steppers.h
#undef CLODE_STEPPERS_H
#undef CLODE_FIXED_STEP
#undef CLODE_ADAPTIVE_STEP
#undef CLODE_STOCHASTIC
#undef CLODE_MAX_ATTEMPTS
#undef N_PAR1
#undef N_AUX1
#undef N_WIENER1
#undef CLODE_ERR_ORDER
#undef CLODE_ROOT_N
#undef CLODE_MAX_SHRINK
#undef CLODE_MAX_GROW
#undef CLODE_EXPON
// clode_observers.h
#undef CLODE_OBSERVERS_H
#undef CLODE_TWO_PASS
#undef CLODE_BASIC_NV
#undef CLODE_BASIC_NA
#undef CLODE_BASIC_VAR
// clode_kernels.h
#undef CLODE_KERNELS_H
#undef CLODE_STEP
