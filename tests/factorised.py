"""FP64 numpy restatement of the engine's factorised scheme (test only).

The HIP engine does not evaluate the reference's 45 stress terms as 45 3-D
inverse transforms; it uses 1-D derivatives of the fields {A1 v, v, A3 v,
A1 dv, dv} and 1-D velocity-step derivatives (see qvti_pencils.h).  This module
restates that factorisation in FP64 with numpy's FFT so that a test can check
it against the literal oracle (63 3-D transforms per step) to rounding: the two
share only the 1-D tables, which are pinned by closed forms elsewhere.
"""
import numpy as np

import oracle as O

PI=3.1415926
def axis_syms(n,d):
    k,dp,dm=O.axis_tables(n,d,'f64')
    dp=dp.copy(); dm=dm.copy()
    if n%2==0: dp[n//2]=dp[n//2].real; dm[n//2]=dm[n//2].real
    return k,dp,dm
def d1(f,sym,axis):
    F=np.fft.fft(f,axis=axis); sh=[1,1,1]; sh[axis]=-1
    return np.fft.ifft(F*sym.reshape(sh),axis=axis).real
def pml(n,p,d,f0):
    a,b=O.pml_profile(n,p,d,f0,'f64'); ah=a.copy(); bh=b.copy()
    ah[:-1]=0.5*(a[:-1]+a[1:]); bh[:-1]=0.5*(b[:-1]+b[1:]); return a,b,ah,bh
def run(shot, m, vti=False, kspace=None):
    nx,ny,nz,P,nt=shot['nx'],shot['ny'],shot['nz'],shot['pml'],shot['nt']
    dt=np.float64(np.float32(shot['dt'])); N=nx*ny*nz
    kx,dpx,dmx=axis_syms(nx,shot['dx']); ky,dpy,dmy=axis_syms(ny,shot['dy']); kz,dpz,dmz=axis_syms(nz,shot['dz'])
    K2=kx[:,None,None]**2+ky[None,:,None]**2+kz[None,None,:]**2
    A1=np.where(K2>0,1/np.sqrt(np.where(K2>0,K2,1)),0); A3=np.sqrt(K2)
    # opt-in k-space factor F(|k|) (the oracle's, pinned by closed forms):
    # F multiplies A1 and A3, and A2 = F becomes a 3-D operator
    F=None
    if kspace:
        F=O.kspace_factor(np.sqrt(K2), np.float32(shot['dt']), kspace.get('vref',0.0), kspace.get('fcut',0.0),
                          kspace.get('vmax',0.0), kspace.get('taper',0.0))
        A1=A1*F; A3=A3*F
    def q(rho,v,g):
        return np.array([O.q_coeffs(r,vv,gg,float(np.float32(shot['w0']))) for r,vv,gg in zip(rho.ravel(),v.ravel(),g.ravel())]).T.reshape(5,nx,ny,nz)
    rho=m['rho'].astype(np.float64)
    cP=q(rho,m['vp'],m['gp']); cS=q(rho,m['vs'],m['gs'])
    if vti:
        vh=m['vp']*np.sqrt(1+2*m['eps'].astype(np.float64)); cH=q(rho,vh,m['gp'])
        c33=rho*m['vp'].astype(np.float64)**2; c44=rho*m['vs'].astype(np.float64)**2
        vd=np.sqrt((np.sqrt(np.maximum((c33-c44)*(c33*(1+2*m['delta'].astype(np.float64))-c44),0))+c44)/rho); cD=q(rho,vd,m['gp'])
    else: cH=cP; cD=cP
    Sb=np.zeros_like(cS)
    Sb[:,:-1,:-1,:-1]=0.125*(cS[:,:-1,:-1,:-1]+cS[:,:-1,:-1,1:]+cS[:,:-1,1:,:-1]+cS[:,:-1,1:,1:]+cS[:,1:,:-1,:-1]+cS[:,1:,:-1,1:]+cS[:,1:,1:,:-1]+cS[:,1:,1:,1:])
    bx=np.zeros((nx,ny,nz)); by=bx.copy(); bz=bx.copy()
    bx[:-1]=dt/(0.5*(rho[:-1]+rho[1:])); by[:,:-1]=dt/(0.5*(rho[:,:-1]+rho[:,1:])); bz[:,:,:-1]=dt/(0.5*(rho[:,:,:-1]+rho[:,:,1:]))
    px=pml(nx,P,shot['dx'],shot['f0']); py=pml(ny,P,shot['dy'],shot['f0']); pz=pml(nz,P,shot['dz'],shot['f0'])
    B=lambda t,ax: t.reshape([-1 if i==ax else 1 for i in range(3)])
    dom=np.zeros((nx,ny,nz),bool); dom[:-1,:-1,:-1]=True
    v=np.zeros((3,nx,ny,nz)); s=np.zeros((6,nx,ny,nz))  # pxx pyy pzz pxy pxz pyz
    phv=np.zeros((9,nx,ny,nz)); phs=np.zeros((14,nx,ny,nz))
    src=(shot['sx'],shot['sy'],shot['sz']); kind=shot['src_kind']
    recs=np.zeros((3,nt,len(shot['rx'])))
    rik=shot['rik'].astype(np.float64); sprev=0.0; dt=np.float64(np.float32(shot['dt']))
    def upd(ph,a,b,d): ph[...] = np.where(dom, ph+dt*(-a*ph-b*d), ph); return ph
    for it in range(nt):
        pxx,pyy,pzz,pxy,pxz,pyz=s
        # velocity: x
        ax_=[d1(pxx,dmx,0),d1(pxy,dpx,0),d1(pxz,dpx,0)]
        ay_=[d1(pxy,dpy,1),d1(pyy,dmy,1),d1(pyz,dpy,1)]
        az_=[d1(pxz,dpz,2),d1(pyz,dpz,2),d1(pzz,dmz,2)]
        hal=[[1,0,0],[0,1,0],[0,0,1]]
        vold=v.copy()
        for c in range(3):
            tot=0
            for a,(dd,pp) in enumerate(zip([ax_[c],ay_[c],az_[c]],[px,py,pz])):
                A_=B(pp[2] if hal[c][a] else pp[0],a); B_=B(pp[3] if hal[c][a] else pp[1],a)
                ph=upd(phv[3*c+a],A_,B_,dd); tot=tot+dd+ph
            v[c]=np.where(dom, v[c]+[bx,by,bz][c]*tot, v[c])
        w=(v-vold)/dt
        if kind<3: w[kind][src]+=sprev/dt
        # H fields
        H1v=[np.fft.ifftn(A1*np.fft.fftn(v[c])).real for c in range(3)]
        H1w=[np.fft.ifftn(A1*np.fft.fftn(w[c])).real for c in range(3)]
        H3v=[np.fft.ifftn(A3*np.fft.fftn(v[c])).real for c in range(3)]
        if F is None: fam=[[H1v[c],v[c],H3v[c],H1w[c],w[c]] for c in range(3)]
        else:
            A2v=[np.fft.ifftn(F*np.fft.fftn(v[c])).real for c in range(3)]
            A2w=[np.fft.ifftn(F*np.fft.fftn(w[c])).real for c in range(3)]
            fam=[[H1v[c],A2v[c],H3v[c],H1w[c],A2w[c]] for c in range(3)]
        S=lambda C,fl,sym,ax: sum(C[j]*d1(fl[j],sym,ax) for j in range(5))
        c12=cH-2*cS; c13d=cD-2*cS
        s11x=S(cH,fam[0],dpx,0); s12x=S(c12,fam[0],dpx,0); s13x=S(c13d,fam[0],dpx,0)
        s12y=S(c12,fam[1],dpy,1); s11y=S(cH,fam[1],dpy,1); s13y=S(c13d,fam[1],dpy,1)
        s13z=S(c13d,fam[2],dpz,2); s33z=S(cP,fam[2],dpz,2)
        sb_xz=S(Sb,fam[0],dmz,2); sb_zx=S(Sb,fam[2],dmx,0); sb_xy=S(Sb,fam[0],dmy,1); sb_yx=S(Sb,fam[1],dmx,0)
        sb_yz=S(Sb,fam[1],dmz,2); sb_zy=S(Sb,fam[2],dmy,1)
        X=lambda t: B(t,0); Y=lambda t: B(t,1); Z=lambda t: B(t,2)
        f1=upd(phs[0],X(px[0]),X(px[1]),s11x); f2=upd(phs[1],X(px[0]),X(px[1]),s12x); f3=upd(phs[2],X(px[0]),X(px[1]),s13x)
        f4=upd(phs[3],X(px[2]),X(px[3]),sb_zx); f5=upd(phs[4],X(px[2]),X(px[3]),sb_yx)
        g1=upd(phs[5],Y(py[0]),Y(py[1]),s12y); g2=upd(phs[6],Y(py[0]),Y(py[1]),s11y); g3=upd(phs[7],Y(py[0]),Y(py[1]),s13y)
        g4=upd(phs[8],Y(py[2]),Y(py[3]),sb_xy); g5=upd(phs[9],Y(py[2]),Y(py[3]),sb_zy)
        h1=upd(phs[10],Z(pz[0]),Z(pz[1]),s13z); h2=upd(phs[11],Z(pz[0]),Z(pz[1]),s33z)
        h3=upd(phs[12],Z(pz[2]),Z(pz[3]),sb_xz); h4=upd(phs[13],Z(pz[2]),Z(pz[3]),sb_yz)
        inc=[s11x+s12y+s13z+f1+g1+h1, s12x+s11y+s13z+f2+g2+h1, s13x+s13y+s33z+f3+g3+h2,
             sb_xy+g4+sb_yx+f5, sb_xz+h3+sb_zx+f4, sb_yz+h4+sb_zy+g5]
        for k in range(6): s[k]=np.where(dom, s[k]+dt*inc[k], s[k])
        if kind<3: v[kind][src]+=rik[it]; sprev=rik[it]
        else:
            for k in range(3): s[k][src]+=rik[it]
        for r,(a,b_,c_) in enumerate(zip(shot['rx'],shot['ry'],shot['rz'])): recs[:,it,r]=v[:,a,b_,c_]
    return recs, np.concatenate([v,s[[0,1,2,3,4,5]]])
