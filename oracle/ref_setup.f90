!*****
!***** Golden-vector driver for the reference's setup numerics.
!*****
! Test infrastructure only.  Built by oracle/Makefile (target "ref") against
! the reference's own MainVar3D.f90 and Subroutineso.f90, compiled where they
! lie; it calls the reference routines penta, fcint, fcbcm, sbcco, mtrxi,
! movef/pitch and indx3 with module state set as src/Main3D.f90:60-63,
! 71-73, 295-332, 443-459 sets it, and prints every result as
! "name count v1 v2 ..." with 17 significant digits.

 program ref_setup

 use mainvar3d
 use subroutineso
 implicit none

 real(kind=nr) :: out5(5), fkv(4), fkb(4), tv(6), dkov(3), dkv(3)
 real(kind=nr), allocatable :: wxu(:,:), wxl(:,:)
 integer(kind=ni) :: q, r, nlen, nsi, nei, nlens(4)

    abc(:,0)=(/a01,a02,a03,a04,a05,a06/)
    abc(:,1)=(/a10,a12,a13,a14,a15,a16/)
    abc(:,2)=(/a20,a21,a23,a24,a25,a26/)
    albed(-2:2,0,0)=(/zero,zero,one,alpha01,beta02/)
    albed(-2:2,1,0)=(/zero,alpha10,one,alpha12,beta13/)
    albed(-2:2,2,0)=(/beta20,alpha21,one,alpha23,beta24/)
    albed(-2:2,0,1)=(/zero,zero,one,alpha,beta/)
    albed(-2:2,1,1)=(/zero,alpha,one,alpha,beta/)
    albed(-2:2,2,1)=(/beta,alpha,one,alpha,beta/)

!===== fcint at several cut-offs (fltr = 1/2 as used by the solver)
    fkv=(/0.85_nr,0.7_nr,0.9_nr,0.6_nr/)
 do q=1,4
    call fcint(pi*fkv(q),half,out5(1),out5(2),out5(3),out5(4),out5(5))
    call emit('fcint',out5,5)
 end do

!===== fcbcm for (fltk, fltkbc) pairs
    fkb=(/0.85_nr,0.7_nr,0.9_nr,0.6_nr/)
 do q=1,4
    call fcbcm(pi*fkv(q),pi*fkb(5-q),albef(:,:,0),fam(:),fbm(:),fcm(:))
    call emit('fcbcm_albef',reshape(albef(:,:,0),(/15/)),15)
    call emit('fcbcm_fam',fam,3)
    call emit('fcbcm_fbm',fbm,3)
    call emit('fcbcm_fcm',fcm,3)
 end do

!===== sbcco closure coefficients (fltk = 0.85 pi and 0.7 pi)
 do q=1,2
    fltk=pi*fkv(q)
    pbco(:,:,:)=0; pbci(:,:,:)=0; call sbcco
    do nt=0,1; do j=0,1; ii=lmd+nt*(lmf-lmd)
       pbcot(j,nt)=sum(pbco(0:ii,j,nt))
    end do; end do
    call emit('sbcco_pbci',reshape(pbci,(/2*2*(lmp+1)/)),2*2*(lmp+1))
    call emit('sbcco_pbco',reshape(pbco,(/2*2*(lmp+1)/)),2*2*(lmp+1))
    call emit('sbcco_pbcot',reshape(pbcot,(/4/)),4)
 end do

!===== penta for deriv and filter matrices, all end-type combinations
    fltk=pi*0.85_nr; fltkbc=pi*0.85_nr
    call fcbcm(fltk,fltkbc,albef(:,:,0),fam(:),fbm(:),fcm(:))
    call fcint(fltk,half,alphf,betf,fa,fb,fc)
    albef(-2:2,0,1)=(/zero,zero,one,alphf,betf/)
    albef(-2:2,1,1)=(/zero,alphf,one,alphf,betf/)
    albef(-2:2,2,1)=(/betf,alphf,one,alphf,betf/)
    nlens=(/7,8,13,40/)
 do q=1,4; nlen=nlens(q)
    lim=nlen-1
    allocate(wxu(0:lim,3),wxl(0:lim,2))
    do nsi=0,1; do nei=0,1
       call penta(wxu,wxl,albed(:,:,nsi),albed(:,:,nei),alpha,beta,0,lim)
       call emit('penta_d_xu',reshape(wxu,(/3*nlen/)),3*nlen)
       call emit('penta_d_xl',reshape(wxl,(/2*nlen/)),2*nlen)
       call penta(wxu,wxl,albef(:,:,nsi),albef(:,:,nei),alphf,betf,0,lim)
       call emit('penta_f_xu',reshape(wxu,(/3*nlen/)),3*nlen)
       call emit('penta_f_xl',reshape(wxl,(/2*nlen/)),2*nlen)
    end do; end do
    deallocate(wxu,wxl)
 end do

!===== GCBC weights (Main3D.f90:443-459, mbci = 3)
    cbca(:,:)=zero; cbca(1,1:2)=albed(1:2,0,0);
    cbca(2,1:3)=albed(0:2,1,0); cbca(3,1:3)=albed(-1:1,2,0)
    rbci(:)=zero; rbci(1:3)=(/one,albed(-1,1,0),albed(-2,2,0)/)
    call mtrxi(cbca,cbcs,1,mbci); sbci(:)=-matmul(cbcs(:,:),rbci(:))
    call emit('gcbc_cbcs',reshape(cbcs,(/9/)),9)
    call emit('gcbc_ra0',(/ra0/),1)
    fctr=pi/(mbci+1); res=zero
 do i=1,mbci; res=res+one
    sbci(i)=half*sbci(i)*(one+cos(res*fctr))
 end do
    call emit('gcbc_sbci',sbci,3)

!===== moving frame (ramp) and pitching frame
    amach1=0.2819077862_nr; amach2=0.102606043_nr; amach3=0
    poo=one/gam; rhooo=one; aoo=sqrt(gam*poo/rhooo)
    amachoo=sqrt(amach1*amach1+amach2*amach2+amach3*amach3)
    uoo(1)=amach1*aoo; uoo(2)=amach2*aoo; uoo(3)=amach3*aoo
    timf=15.0_nr; nsmf=0
    tv=(/0.0_nr,0.37_nr,3.1_nr,14.9_nr,15.2_nr,40.0_nr/)
    dkov=(/0.0_nr,0.01_nr,0.02_nr/); dkv=(/0.0_nr,0.013_nr,0.04_nr/)
 do r=1,2
    if(r==1) then; tp=0; tps=0; aoa0=20; aoa1=10
    else; tp=12.0_nr; tps=2.0_nr; aoa0=20; aoa1=10; end if
    do q=1,6; do i=1,3
       timo=tv(q); dtko=dkov(i); dtk=dkv(i); aoa=aoa0
       call movef(dtko,dtk)
       call emit('movef_umf',umf,3)
       call emit('movef_dudtmf',dudtmf,3)
       call emit('movef_aoa',(/aoa/),1)
    end do; end do
 end do

!===== index map
    lxi=5; let=6; lze=7
 do nn=1,3
    call emit('indx3',(/real(indx3(1,2,3,nn),nr),real(indx3(4,0,2,nn),nr),&
                        real(indx3(0,5,1,nn),nr)/),3)
 end do

 contains

 subroutine emit(name,v,nv)
 character(len=*),intent(in) :: name
 integer(kind=ni),intent(in) :: nv
 real(kind=nr),intent(in) :: v(nv)
 integer(kind=ni) :: s
    write(*,'(a,1x,i0)',advance='no') name,nv
 do s=1,nv
    write(*,'(1x,es25.17)',advance='no') v(s)
 end do
    write(*,*)
 end subroutine emit

 end program ref_setup
