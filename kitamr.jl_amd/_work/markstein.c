#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
static uint64_t s=88172645463325252ull;
static uint64_t xr(){s^=s<<13;s^=s>>7;s^=s<<17;return s;}
static double rnd(int lo,int hi){ // random double with exponent in [lo,hi]
  uint64_t m=xr()&((1ull<<52)-1); int e=lo+(int)(xr()%(uint64_t)(hi-lo+1));
  double x=ldexp(1.0+ldexp((double)m,-52),e); return (xr()&1)?-x:x;}
int main(){
  long bad=0,n=0;
  for(long t=0;t<200000000;++t){
    double a=rnd(-120,40), b=rnd(-40,10);
    if(t%7==0) a=(double)(xr()%1000)*0.001 - 0.5;
    double r=1.0/b; double q=a*r; double e=fma(-q,b,a); double q2=fma(e,r,q);
    double ref=a/b; ++n;
    if(q2!=ref && !(isnan(q2)&&isnan(ref))){ if(bad<5) printf("a=%a b=%a ref=%a got=%a\n",a,b,ref,q2); ++bad;}
  }
  printf("checked %ld, mismatches %ld\n",n,bad); return 0;}
