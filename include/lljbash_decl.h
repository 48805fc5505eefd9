/* Symbol-naming glue shared by every public header.
 *
 * Replaces include/lljbash_decl.h of lljbash/par_ilu0_cpp (lines 1-21): every
 * exported C name carries the prefix LLJBASH_, and C++ translation units see
 * default member initialisers while C ones see plain struct declarations.
 * The three macros keep the reference's names so that code written against
 * the reference headers compiles unchanged against these.
 */
#ifndef LLJBASH_DECL_H_
#define LLJBASH_DECL_H_

#ifndef LLJBASH_DECL
#define LLJBASH_DECL(name) LLJBASH_##name
#endif

#ifndef LLJBASH_DEFAULT_VALUE
#  ifdef __cplusplus
#    define LLJBASH_DEFAULT_VALUE(v) = v
#  else
#    define LLJBASH_DEFAULT_VALUE(v)
#  endif
#endif

#ifndef LLJBASH_STRUCT_TYPEDEF
#  ifdef __cplusplus
#    define LLJBASH_STRUCT_TYPEDEF(tag)
#  else
#    define LLJBASH_STRUCT_TYPEDEF(tag) typedef struct tag tag
#  endif
#endif

#endif /* LLJBASH_DECL_H_ */
