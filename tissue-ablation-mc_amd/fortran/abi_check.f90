! abi_check.f90 -- the Fortran interoperable types against the C structs:
! prints each pair of sizes and stops with status 1 on any difference.
program abi_check

    use iso_c_binding
    use tam_bind
    implicit none

    type(tam_params)   :: p
    type(tam_material) :: m
    type(tam_mc_stats) :: s
    type(tam_pulse)    :: q
    type(tam_timing)   :: t
    integer(c_int64_t) :: f(0:4), c(0:4)
    integer :: i

    f = [int(c_sizeof(p), c_int64_t), int(c_sizeof(m), c_int64_t), int(c_sizeof(s), c_int64_t), &
         int(c_sizeof(q), c_int64_t), int(c_sizeof(t), c_int64_t)]
    do i = 0, 4
        c(i) = tam_struct_size(int(i, c_int32_t))
        write(*, '(a, i0, a, i0, a, i0)') "struct ", i, ": fortran ", f(i), " c ", c(i)
    end do
    write(*, '(a, i0)') "abi version ", tam_abi_version()
    if (any(f /= c)) error stop 1
end program abi_check
